"""Benchmark of the SPH step on one MI355X.

Workload (default): the mixed batch -- 64 independent scenes, dynamic particle
counts log-uniform in [2k, 60k] (seed 3), viscosity model per scene from
{Newtonian, Cross, power law}, eta0 log-uniform in [0.1, 100] solved by the
Weiler implicit solver, 25% of scenes with an elastic block, dt 5e-4.

  value     particle-steps/s of the batched form: all 64 scenes resident in one
            pool, one launch sequence per tick (a particle-step = one dynamic
            particle advanced one step; boundary particles are not counted)
  plain     the same scenes stepped one at a time, one context each, over the
            same step indices as the batched form (same warm-up, same count)
  latency   p50/p99 scene-step latency: the tick time for the batched form,
            the per-scene step time for the plain form
  roofline  the kernel with the largest share of a profiled tick: algorithmic
            bytes per launch (per-kernel table below, from the logged per-scene
            counts and iteration numbers) over its mean launch duration, from
            HIP events around every 8th launch of it in the timed region
  cpu       the oracle (FP32, the reference's arithmetic, OpenMP over particles)
            on the host cores, on a bounded sample of the same scenes

    python bench.py [--steps K] [--warmup W] [--workload mixed_batch|dam_break|nonnewton_dam|elastic_drop]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from nonnewton_amd import _lib as L  # noqa: E402
from nonnewton_amd import scenes as S  # noqa: E402

METRIC = "particle-steps/s batched FP32 vs plain; p50/p99 scene-step latency, mixed load"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# algorithmic bytes per launch, per scene (4-byte words; per-particle inputs
# counted once, one word per neighbor-list entry walked).  N all particles,
# Nf fluid, E = sum of neighbor counts, Ef = sum over fluid particles.
# ---------------------------------------------------------------------------
def alg_bytes(kernel, N, Nf, Ne, E, Ef, warm=0):
    w = {
        "build_lists": 4 * N + E,               # x in; lists + counts out
        "density_alpha": 8 * N + 4 * Nf + Ef,   # x,V,v,kind in; rho, alpha, rho-dot out
        "div_update": 5 * N + 9 * Nf + Ef,      # x,V,kind; rho-dot, alpha; v in/out
        "div_source": 8 * N + 2 * Nf + Ef,      # x,V,v,kind; rho-dot out
        "press_update": 5 * N + 9 * Nf + Ef,
        "press_source": 8 * N + 3 * Nf + Ef,    # + rho in
        "nonpressure": 14 * N + 2 * Nf + Ef,    # x,V,v,kind, rho_j; acc and v' out
        # the CG's 3-vectors are 3 words (u, w, p, s, x, r, e; M is 6)
        "cg_matvec": 4 * N + 8 * Nf + 2 * Ef,   # x, kind; u, rho, count in; w out; list + pair scalar
        "cg_update": N + 39 * Nf,               # kind; u, w, p, s, x, r, M in; p, s, x, r, u out
        "cg_init": (13 + 3 * warm) * N + (10 + 3 * warm) * Nf + 2 * Ef,  # x,V,v,mu,rho,kind (+e) in,
                                                # x out; count in, M, r0 (+Ae) out; list in, pair scalar out
        "cg_guess": N + (21 + 6 * warm) * Nf,   # kind; v, r0, M (+e, Ae) in; x, r, u out
        "strain": 8 * N + 3 * Nf + Ef,
        "advect": 12 * N,
        "clear_acc": 5 * N,
        "elastic_rot_stress": 23 * Ne,
        "elastic_force": 26 * Ne,
    }.get(kernel)
    return None if w is None else 4 * w


# iteration kernels run once per solver iteration of each active scene
ITER_KERNEL = {"div_update": "iterations_v", "div_source": "iterations_v", "press_update": "iterations_p",
               "cg_update": "iterations_cg"}


def scene_algorithmic_bytes(kernel, sc, st):
    b = alg_bytes(kernel, sc.n, sc.n_fluid, sc.n_elastic, st["num_pairs"], st["num_pairs_fluid"],
                  int(sc.params.warm_start) & 1)
    if b is None:
        return None
    p = sc.params
    weiler = sc.n_fluid > 0 and p.use_viscosity and p.viscosity_method == 1
    if kernel.startswith("cg_") and not weiler:
        return 0
    if kernel == "strain" and not (weiler or p.use_non_newtonian):
        return 0
    if kernel in ITER_KERNEL:
        return b * st[ITER_KERNEL[kernel]]
    if kernel == "press_source":  # one launch before the loop plus one per iteration
        return b * (1 + st["iterations_p"])
    if kernel == "cg_matvec":  # one launch at the start of the solve plus one per iteration
        return b * (1 + st["iterations_cg"])
    if kernel == "div_source":
        return b * st["iterations_v"]
    return b


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (profiles/<kernel>_<workload>_traffic.md,
    counters scaled by a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def load_workload(name):
    if name == "mixed_batch":
        return S.mixed_batch(), "mixed_batch: 64 scenes, dynamic particles log-uniform 2k-60k, seed 3, dt 5e-4"
    sc = S.by_name(name)
    return [sc], f"{name}: {sc.n} particles ({sc.n_dynamic} dynamic), dt {sc.params.dt}"


EVENT_STRIDE = 8  # the dominant kernel is timed on every 8th launch in the timed region


def _admit_all(scs, block):
    pool = L.Pool(200 << 30)
    if block:
        pool.set_block(block)
    slots = [pool.admit(s.params, s.x, s.kind, s.v) for s in scs]
    if any(q is None for q in slots):
        raise RuntimeError("pool budget too small for the workload")
    return pool, slots


def run_batched(scs, steps, warmup, profile=True, block=0, tune=True, kernel=None):
    """The batched form.  Launch shapes are tuned on a separate pool holding the
    same scenes (the tuner advances its scenes), so the timed pool starts from
    the same state whether tuned or not."""
    from nonnewton_amd import autotune
    shapes = {}
    if tune:
        tp, tslots = _admit_all(scs, block)
        for k in range(warmup):
            tp.tick()
        shapes = autotune.tune_pool(tp, tslots, [s.n_fluid for s in scs], use_cache=False)
        tp.close()
        log(f"[bench] launch shapes: {shapes}")
    t0 = time.time()
    pool, slots = _admit_all(scs, block)
    autotune.apply(pool, shapes)
    log(f"[bench] admitted {len(slots)} scenes in {time.time() - t0:.2f} s")
    for k in range(warmup):
        pool.tick()
    breakdown = {}
    dominant = None
    if profile:
        # one more untimed tick with every kernel timed picks the dominant
        # kernel; the timed region then records events around every
        # EVENT_STRIDE-th launch of that kernel only
        pool.profile(True)
        pool.tick()
        breakdown = pool.kernel_times()
        dominant = kernel or max(breakdown.items(), key=lambda kv: kv[1][0])[0]
        pool.profile(True, kernels=[dominant], stride=EVENT_STRIDE)
    lat = []
    raw = []
    t_start = time.perf_counter()
    for k in range(steps):
        t = time.perf_counter()
        pool.tick()
        lat.append(time.perf_counter() - t)
        raw.append(pool.stats_array(slots))
    elapsed = time.perf_counter() - t_start
    names = raw[0].dtype.names if raw else ()
    per_step_stats = [[{f: a[f][k].item() for f in names} for k in range(len(slots))] for a in raw]
    times = {}
    launches = {}
    if profile:
        ms, timed = pool.kernel_times()[dominant]
        launches = pool.kernel_launches()  # every kernel's launches in the timed region
        times = {dominant: (ms, timed, launches[dominant])}
    pool.profile(False)
    pool.close()
    return elapsed, lat, per_step_stats, times, breakdown, shapes, launches


def run_plain(scs, steps, warmup, block=0, shapes=None):
    ctxs = []
    for s in scs:
        c = L.Context(s.params, s.n)
        if block:
            c.set_block(block)
        for k, sh in (shapes or {}).items():  # the batched form's tuned shapes
            c.set_kernel_block(k, sh["block"])
            if sh.get("depth"):
                c.set_kernel_depth(k, sh["depth"])
        c.upload(s.x, s.kind, s.v)
        c.init()
        for k in range(warmup):
            c.step()
        ctxs.append(c)
    lat = []
    its = []
    t_start = time.perf_counter()
    for k in range(steps):
        for c in ctxs:
            t = time.perf_counter()
            st = c.step()
            lat.append(time.perf_counter() - t)
            its.append((st["iterations_v"], st["iterations_p"], st["iterations_cg"]))
    elapsed = time.perf_counter() - t_start
    for c in ctxs:
        c.close()
    return elapsed, lat, np.array(its, dtype=np.float64)


def run_serving(count, rate):
    """Continuous batching over a Poisson request stream (sizes log-uniform 1k-500k,
    50-500 steps, seed 4), 200 GB budget."""
    from nonnewton_amd.serving import ServingLoop

    reqs = S.serving_requests(seed=4, rate=rate, count=count)
    loop = ServingLoop(200 << 30)
    rep = loop.run(reqs)
    loop.close()
    s = rep.summary()
    s["rate_per_s"] = rate
    s["dynamic_particles_requested"] = int(sum(q.scene.n_dynamic for q in reqs))
    s["steps_requested"] = int(sum(q.steps for q in reqs))
    return s


def cpu_baseline(scs, budget_s):
    """The oracle in FP32 without compensated sums (the reference's arithmetic)
    on the host cores: one untimed step, then timed steps, scene after scene,
    until about budget_s seconds of timed CPU work."""
    from oracle import oracle as O

    cores = len(os.sched_getaffinity(0))
    O.set_threads(cores, "f32")
    done_dyn = 0
    steps = 0
    spent = 0.0
    sims = []
    for s in scs:
        sim = O.OracleSim.from_scene(s, precision="f32", compensated=False)
        sim.init()
        sim.step()
        sims.append((s, sim))
    passes = 0
    while spent < budget_s:
        # round robin over every scene, so the sample has the workload's mix
        for s, sim in sims:
            t = time.perf_counter()
            sim.step()
            spent += time.perf_counter() - t
            done_dyn += s.n_dynamic
            steps += 1
        passes += 1
    return {
        "value": done_dyn / spent,
        "unit": "particle-steps/s",
        "cores": int(O.num_threads("f32")),
        "kind": "port",
        "sample": f"oracle FP32 (reference arithmetic, OpenMP over particles): {passes} round-robin pass(es) of one "
                  f"timed step over all {len(sims)} scenes of the workload after one untimed step each; {steps} "
                  f"scene-steps, {done_dyn} particle-steps, {spent:.1f} s of timed CPU work",
    }


def copy_probe(mib):
    """A plain device-to-device copy of known size: each launch reads and writes
    mib MiB, so FETCH_SIZE and WRITE_SIZE of its dispatches (KiB) give the
    counters' scale for HBM bytes.  Uses the HIP runtime directly (hipMemcpy
    device to device runs the runtime's copy kernel)."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mib * (1 << 20)
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(5):
        if hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes, "launches": 5}),
          flush=True)


def surface_block(n, sp=0.02, seed=5):
    """A jittered lattice block of about n particles (aspect 2:1:2, spacing sp)."""
    k = max(2, round((n / 4.0) ** (1.0 / 3.0)))
    dims = (2 * k, k, 2 * k)
    rng = np.random.Generator(np.random.PCG64(seed))
    g = [np.arange(d) * sp for d in dims]
    x = np.stack(np.meshgrid(*g, indexing="ij"), -1).reshape(-1, 3) + 0.1
    x += rng.uniform(-0.3 * sp, 0.3 * sp, size=x.shape)
    return x.astype(np.float32)


def run_surface(n, reps, cpu_particles):
    """Surface reconstruction (SRTool.dual_contouring, SPH/surf_recon/srtool.py:1264-1304)
    of a jittered fluid block at the reference example's ratios (voxel 0.01,
    particle radius 0.02, smooth radius 0.06, spacing 0.02): positions resident
    in device memory, one warm-up run, then `reps` timed runs (each a whole
    reconstruction, host synchronisations included)."""
    import ctypes

    from nonnewton_amd.surface import Surface

    x = surface_block(n)
    kw = dict(voxel_size=0.01, particle_radius=0.02, smooth_radius=0.06, isolevel=0.0)
    hip = ctypes.CDLL("libamdhip64.so.7")
    d = ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(d), ctypes.c_size_t(x.nbytes)) != 0:
        raise RuntimeError("hipMalloc failed")
    if hip.hipMemcpy(d, x.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(x.nbytes), 1) != 0:
        raise RuntimeError("hipMemcpy failed")
    s = Surface(**kw)
    info = s.run_device(d.value, len(x))
    t0 = time.perf_counter()
    for _ in range(reps):
        info = s.run_device(d.value, len(x))
    el = time.perf_counter() - t0
    s.close()
    hip.hipFree(d)
    out = {
        "metric": "surface reconstruction particles/s (SRTool.dual_contouring)",
        "value": round(len(x) * reps / el, 1),
        "unit": "particles/s",
        "ms_per_reconstruction": round(1e3 * el / reps, 3),
        "reps": reps,
        "dtype": "f32",
        "data": "synthetic (seeded jittered block)",
        "config": {"workload": f"jittered block, {len(x)} particles, spacing 0.02", **kw,
                   "grid_vertices": int(np.prod(info["dims"])), "written_blocks": info["active_blocks"],
                   "mesh_vertices": info["num_vertices"], "mesh_triangles": info["num_triangles"]},
    }
    if cpu_particles > 0:
        from oracle import surface as O

        xs = surface_block(cpu_particles)
        t0 = time.perf_counter()
        O.reconstruct(xs, **kw)
        ce = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(len(xs) / ce, 1), "unit": "particles/s", "cores": 1, "kind": "port",
                               "sample": f"one reconstruction of a {len(xs)}-particle block of the same kind "
                                         f"({ce:.1f} s)"}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=0, help="timed steps of the plain form (0 = --steps)")
    ap.add_argument("--workload", default="mixed_batch")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (no autotuning)")
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--roofline-kernel", default=None, metavar="KERNEL",
                    help="report the roofline of KERNEL instead of the one that dominates the probe tick")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 5 times (calibrates the rocprofv3 FETCH_SIZE and "
                         "WRITE_SIZE counters against a kernel of known traffic) and exit")
    ap.add_argument("--surface", type=int, default=0, metavar="N",
                    help="only measure the surface reconstruction of an N-particle block and exit")
    ap.add_argument("--surface-reps", type=int, default=5)
    ap.add_argument("--surface-cpu", type=int, default=16000, metavar="N",
                    help="particles of the oracle's CPU sample for --surface (0 = none)")
    args = ap.parse_args()
    if args.surface:
        run_surface(args.surface, args.surface_reps, args.surface_cpu)
        return
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the SPH scenes are independent and shard across GPUs without exchange; "
              "run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    scs, workload = load_workload(args.workload)
    n_dyn = sum(s.n_dynamic for s in scs)
    n_all = sum(s.n for s in scs)
    log(f"[bench] workload {workload}: {len(scs)} scenes, {n_all} particles, {n_dyn} dynamic")

    elapsed, lat, stats, ktimes, breakdown, shapes, launch_counts = run_batched(
        scs, args.steps, args.warmup, block=args.block, tune=not args.no_tune, kernel=args.roofline_kernel)
    value = n_dyn * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    log(f"[bench] batched: {ms_per_step:.3f} ms/tick, {value:.4g} particle-steps/s")

    plain_value = None
    plain_lat = []
    if not args.no_plain:
        # the same step indices as the batched form's timed ticks (same warm-up,
        # same count), so both forms do the same solver work
        plain_steps = args.plain_steps or args.steps
        pe, plain_lat, plain_its = run_plain(scs, plain_steps, args.warmup, block=args.block, shapes=shapes)
        plain_value = n_dyn * plain_steps / pe
        log(f"[bench] plain: {1e3 * pe / plain_steps:.3f} ms per round of {len(scs)} scenes, "
            f"{plain_value:.4g} particle-steps/s")

    # roofline of the dominant kernel
    roof = None
    if ktimes:
        name, (ms, timed, launches) = max(ktimes.items(), key=lambda kv: kv[1][0])
        total_bytes = 0.0
        ok = True
        for step_stats in stats:
            for sc, st in zip(scs, step_stats):
                b = scene_algorithmic_bytes(name, sc, st)
                if b is None:
                    ok = False
                    break
                total_bytes += b
        if ok and launches and timed:
            avg_s = ms * 1e-3 / timed          # mean duration of the timed (every EVENT_STRIDE-th) launches
            per_launch = total_bytes / launches  # algorithmic bytes per launch over all launches
            achieved = per_launch / avg_s / 1e9
            roof = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(name, args.workload),
                    "launches": launches, "timed_launches": timed, "avg_launch_us": round(1e6 * avg_s, 2),
                    "bytes_per_launch": round(per_launch),
                    "share_of_timed_region": round(avg_s * launches / elapsed, 3)}
        # algorithmic bytes per launch of the other CG kernels over the same ticks
        # (the rocprofv3 traffic summaries compare their counters with these)
        per_kernel = {}
        for k in ("cg_matvec", "cg_update"):
            n_l = launch_counts.get(k, 0)
            tb = sum(scene_algorithmic_bytes(k, sc, st) or 0 for ss in stats for sc, st in zip(scs, ss))
            if n_l:
                per_kernel[k] = round(tb / n_l)
        if roof is not None:
            roof["alg_bytes_per_launch"] = per_kernel
        log("[bench] kernel times of one tick (ms, launches): " +
            ", ".join(f"{k}={v[0]:.2f}/{v[1]}" for k, v in sorted(breakdown.items(), key=lambda kv: -kv[1][0])
                      if v[1]))

    it = np.array([[st["iterations_v"], st["iterations_p"], st["iterations_cg"]] for ss in stats for st in ss])
    cfg = {
        "workload": workload,
        "scenes": len(scs),
        "particles": n_all,
        "dynamic_particles": n_dyn,
        "batched_ms_per_tick": round(ms_per_step, 3),
        "p50_scene_step_ms_batched": round(1e3 * float(np.percentile(lat, 50)), 3),
        "p99_scene_step_ms_batched": round(1e3 * float(np.percentile(lat, 99)), 3),
        "mean_iterations_v_p_cg": [round(float(x), 2) for x in it.mean(axis=0)],
        "launch_shapes": shapes,
    }
    if plain_value is not None:
        cfg["plain_particle_steps_per_s"] = round(plain_value, 1)
        cfg["batched_over_plain"] = round(value / plain_value, 3)
        cfg["p50_scene_step_ms_plain"] = round(1e3 * float(np.percentile(plain_lat, 50)), 3)
        cfg["p99_scene_step_ms_plain"] = round(1e3 * float(np.percentile(plain_lat, 99)), 3)
        cfg["plain_steps_timed"] = plain_steps
        cfg["mean_iterations_v_p_cg_plain"] = [round(float(x), 2) for x in plain_its.mean(axis=0)]

    if args.serve_requests > 0 and args.workload == "mixed_batch":
        cfg["serving"] = run_serving(args.serve_requests, args.serve_rate)
        log(f"[bench] serving: {cfg['serving']}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scs, args.cpu_seconds)
        log(f"[bench] cpu baseline: {cpu['value']:.4g} particle-steps/s on {cpu['cores']} cores")

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded lattice scenes, no external data)",
        "config": cfg,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
