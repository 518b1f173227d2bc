"""Benchmark of the SPH step on one MI355X.

Workload (default): the mixed batch -- 64 independent scenes, dynamic particle
counts log-uniform in [2k, 60k] (seed 3), viscosity model per scene from
{Newtonian, Cross, power law}, eta0 log-uniform in [0.1, 100] solved by the
Weiler implicit solver, 25% of scenes with an elastic block, dt 5e-4.

  value     particle-steps/s of the batched form: all 64 scenes resident,
            spread over --lanes pools ("lanes", default 3) that are stepped
            concurrently, each by its own host thread on its own stream with
            one launch sequence per tick (a particle-step = one dynamic
            particle advanced one step; boundary particles are not counted);
            the same scenes in one pool are timed too (one_lane_*)
  plain     the same scenes stepped one at a time, one context each, over the
            same step indices as the batched form (same warm-up, same count)
  control   one scene pooled against the same scene in a context (about 1)
  latency   p50/p99 scene-step latency: the lanes' tick times for the batched
            form, the per-scene step time for the plain form
  roofline  the kernel with the largest share of a probe tick (on a separate
            pool): algorithmic bytes per launch (sweep table below, from the
            logged per-scene counts and iteration numbers) over its mean
            launch duration, from HIP events carried by every 8th launch of it
            in the one-lane form's timed region, less the launch and event
            latency an empty kernel shows when bracketed the same way; also a
            whole CG iteration, and the whole tick of the lanes and of the
            one-lane form
  serving   continuous batching of the Poisson request stream (1k-500k
            particles, 50-500 steps) at 5, 20 and 80 requests/s
  cpu       the oracle (FP32, the reference's arithmetic, OpenMP over particles)
            on the host cores, on a bounded sample of the same scenes, and the
            oracle as the checker: neighbor lists and fields of two scenes,
            the dam break's density-error drift against the FP64 oracle

The one-lane form's timed ticks (or, with --trace-region lanes, the lanes')
are bracketed by roctxProfilerResume/Pause, so
`rocprofv3 --selected-regions --kernel-trace --stats -- python bench.py ...`
profiles exactly them.

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
# Algorithmic bytes: 4 bytes per word; per-particle inputs counted once for all
# N particles a sweep touches, plus the words it writes, plus one word per
# neighbor-list entry it walks (re-reads of neighbor data are free).  N all
# particles of a scene, Nf fluid, Ne elastic, Nd dynamic solids (elastic and
# rigid), Ndyn dynamic (fluid + Nd); E the list entries of the step's search
# (fluid and dynamic-solid lists), Ef those of fluid particles, Ed those of
# dynamic solids (also taken for the elastic rest lists).
# ---------------------------------------------------------------------------
def sweep_words(sweep, N, Nf, Ne, Nd, E, Ef, Ed):
    return {
        "search": 4 * N + E,                 # x in; cell offsets and lists out
        "density_alpha": 5 * N + 2 * Nf + Ef,
        "density_change": 7 * N + Nf + Ef,   # + its FP64 reduction
        "div_update": 5 * N + 8 * Nf + Ef,
        "nonpressure": 7 * N + 5 * Nf + Ef,  # viscosity, surface tension, advect of v
        "density_adv": 7 * N + 2 * Nf + Ef,  # + its FP64 reduction
        "press_update": 5 * N + 8 * Nf + Ef,
        "advect": 12 * (Nf + Nd),
        "boundary_volume": 4 * N + Nd + Ed,
        "elastic_rot_stress": 8 * Ne + 15 * Ne + Ed,
        "elastic_force": 23 * Ne + 3 * Ne + Ed,
        # one Weiler CG iteration: w = A u over the lists (x, u, rho, kind of
        # fluid particles in, w out: 11 Nf + Ef) and the vector updates with
        # the block preconditioner (37 Nf) -- 9 Nf + 3 Nf + Ef + 36 Nf in all
        "cg_matvec": 11 * Nf + Ef,
        "cg_update": 37 * Nf,
        "strain": 8 * N + Nf + Ef,
    }[sweep]


def scene_counts(sc, st):
    nd = sc.n_elastic + sc.n_rigid
    E, Ef = st["num_pairs"], st["num_pairs_fluid"]
    return dict(N=sc.n, Nf=sc.n_fluid, Ne=sc.n_elastic, Nd=nd, E=E, Ef=Ef, Ed=E - Ef)


def _weiler(sc):
    p = sc.params
    return sc.n_fluid > 0 and p.use_viscosity and p.viscosity_method == 1


def kernel_algorithmic_bytes(kernel, sc, st):
    """Algorithmic bytes of `kernel`'s launches for scene `sc` in one step with
    statistics `st`; None for kernels without a table entry."""
    c = scene_counts(sc, st)
    p = sc.params
    Iv, Ip, Icg = st["iterations_v"], st["iterations_p"], st["iterations_cg"]
    fluid = sc.n_fluid > 0

    def w(name):
        return 4 * sweep_words(name, **c)

    if kernel == "build_lists":
        return w("search")
    if kernel == "density_alpha":  # density and alpha, and the first density change (fused)
        return (w("density_alpha") + (w("density_change") if p.enable_divergence_solver else 0)) if fluid else 0
    if kernel == "div_source":
        return Iv * w("density_change") if fluid else 0
    if kernel == "div_update":
        return Iv * w("div_update") if fluid else 0
    if kernel == "nonpressure":
        return w("nonpressure") if fluid else 0
    if kernel == "press_source":
        return (1 + Ip) * w("density_adv") if fluid else 0
    if kernel == "press_update":
        return Ip * w("press_update") if fluid else 0
    if kernel == "advect":
        return w("advect")
    if kernel == "bvol":
        return w("boundary_volume") if c["Nd"] else 0
    if kernel in ("elastic_rot_stress", "elastic_force"):
        return w(kernel) if c["Ne"] else 0
    if kernel == "strain":
        return w("strain") if (_weiler(sc) or p.use_non_newtonian) else 0
    if kernel == "cg_matvec":  # the solve's first product, then one per iteration
        return (1 + Icg) * w("cg_matvec") if _weiler(sc) else 0
    if kernel == "cg_update":
        return Icg * w("cg_update") if _weiler(sc) else 0
    return None


def step_algorithmic_bytes(sc, st):
    """The whole step: search + density/alpha + (1 + Iv) density changes + Iv
    divergence updates + non-pressure + (1 + Ip) predicted densities + Ip
    pressure updates + advect, with the boundary-volume, elastic, strain and
    Weiler terms (Icg iterations) where the scene has them."""
    c = scene_counts(sc, st)
    p = sc.params
    Iv, Ip, Icg = st["iterations_v"], st["iterations_p"], st["iterations_cg"]

    def w(name):
        return 4 * sweep_words(name, **c)

    b = w("search") + w("advect")
    if sc.n_fluid:
        b += w("density_alpha") + w("nonpressure")
        if p.enable_divergence_solver:
            b += (1 + Iv) * w("density_change") + Iv * w("div_update")
        b += (1 + Ip) * w("density_adv") + Ip * w("press_update")
        if _weiler(sc) or p.use_non_newtonian:
            b += w("strain")
        if _weiler(sc):
            b += Icg * (w("cg_matvec") + w("cg_update"))
    if c["Nd"]:
        b += w("boundary_volume")
    if c["Ne"]:
        b += w("elastic_rot_stress") + w("elastic_force")
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


EVENT_STRIDE = 8  # the timed kernels carry HIP events on every 8th launch in the timed region


class TimedRegion:
    """Brackets the timed ticks with roctxProfilerResume / roctxProfilerPause,
    so that `rocprofv3 --selected-regions` records exactly the launches the
    line reports on (without a profiler the calls do nothing)."""

    def __init__(self):
        import ctypes

        self.lib = None
        for name in ("librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so"):
            try:
                self.lib = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if self.lib is not None:
            for fn in (self.lib.roctxProfilerResume, self.lib.roctxProfilerPause):
                fn.argtypes = [ctypes.c_uint64]
                fn.restype = ctypes.c_int

    def __enter__(self):
        if self.lib is not None:
            self.lib.roctxProfilerResume(0)
        return self

    def __exit__(self, *exc):
        if self.lib is not None:
            self.lib.roctxProfilerPause(0)
        return False


def _admit_all(scs, block):
    pool = L.Pool(200 << 30)
    if block:
        pool.set_block(block)
    slots = [pool.admit(s.params, s.x, s.kind, s.v) for s in scs]
    if any(q is None for q in slots):
        raise RuntimeError("pool budget too small for the workload")
    return pool, slots


def run_batched(scs, steps, warmup, profile=True, block=0, tune=True, kernel=None, region=None):
    """The batched form.  Launch shapes are tuned, and one tick with every
    kernel timed picks the dominant kernel, on a separate pool holding the same
    scenes; the timed pool then runs `warmup` ticks and the `steps` timed ticks,
    so it times the same step indices as the plain form.  In the timed region
    the dominant kernel and the CG kernels carry HIP events on every
    EVENT_STRIDE-th launch."""
    from nonnewton_amd import autotune
    shapes = {}
    breakdown = {}
    tp, tslots = _admit_all(scs, block)
    for k in range(warmup):
        tp.tick()
    if tune:
        shapes = autotune.tune_pool(tp, tslots, [s.n_fluid for s in scs], use_cache=False)
        log(f"[bench] launch shapes: {shapes}")
    if profile:
        tp.profile(True)
        tp.tick()
        breakdown = tp.kernel_times()
        tp.profile(False)
    tp.close()
    t0 = time.time()
    pool, slots = _admit_all(scs, block)
    autotune.apply(pool, shapes)
    log(f"[bench] admitted {len(slots)} scenes in {time.time() - t0:.2f} s")
    for k in range(warmup):
        pool.tick()
    dominant = None
    timed_kernels = []
    if profile:
        dominant = kernel or max(breakdown.items(), key=lambda kv: kv[1][0])[0]
        timed_kernels = [dominant] + [k for k in ("cg_matvec", "cg_update") if k != dominant]
        pool.profile(True, kernels=timed_kernels, stride=EVENT_STRIDE)
    lat = []
    raw = []
    with (region or TimedRegion()):
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
        kt = pool.kernel_times()
        launches = pool.kernel_launches()  # every kernel's launches in the timed region
        times = {k: (kt[k][0], kt[k][1], launches[k]) for k in timed_kernels if launches.get(k)}
    pool.profile(False)
    pool.close()
    return dict(elapsed=elapsed, lat=lat, stats=per_step_stats, times=times, dominant=dominant,
                breakdown=breakdown, shapes=shapes, launches=launches)


def run_lanes(scs, steps, warmup, lanes, block=0, shapes=None, region=None):
    """The batched form in `lanes` lanes (LanePool): the scenes admitted largest
    first, each to the least loaded lane (particles), every lane stepped by its
    own host thread on its own stream with no barrier between lanes.  `warmup`
    ticks of every lane, then `steps` timed ticks of every lane: the timed
    region ends when the last lane has done its `steps` ticks, so every scene
    advances exactly `steps` steps in it, the same step indices as the plain
    form."""
    from nonnewton_amd.lanes import LanePool

    lp = LanePool(200 << 30, lanes=lanes, shapes=shapes, block=block)
    order = sorted(range(len(scs)), key=lambda k: (-scs[k].n, k))
    keys = {}
    for k in order:
        s = scs[k]
        key = lp.admit(s.params, s.x, s.kind, s.v, n_elastic=s.n_elastic)
        if key is None:
            raise RuntimeError("pool budget too small for the workload")
        keys[k] = key
    slots = [[keys[k][1] for k in sorted(keys) if keys[k][0] == l] for l in range(lanes)]
    scene_of = [[k for k in sorted(keys) if keys[k][0] == l] for l in range(lanes)]
    lp.run(warmup, [[] for _ in range(lanes)])
    with (region or _NoRegion()):
        elapsed, out = lp.run(steps, slots)
    lp.close()
    lat = [t for lane in out for t in lane[0]]
    # per tick k: every scene's statistics in workload order
    names = next(o[1][0].dtype.names for o in out if o[1])
    stats = []
    for k in range(steps):
        row = [None] * len(scs)
        for l in range(lanes):
            if not scene_of[l]:
                continue
            a = out[l][1][k]
            for q, sk in enumerate(scene_of[l]):
                row[sk] = {f: a[f][q].item() for f in names}
        stats.append(row)
    return dict(elapsed=elapsed, lat=lat, stats=stats, lanes=lanes,
                scenes_per_lane=[len(x) for x in scene_of],
                particles_per_lane=[int(sum(scs[k].n for k in x)) for x in scene_of],
                lane_ms_per_tick=[round(1e3 * sum(o[0]) / steps, 3) for o in out])


class _NoRegion:
    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


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


def one_scene_control(sc, steps, warmup, shapes):
    """One scene in a pool against the same scene in a context, same steps: the
    two paths run the same kernels, so the ratio is about 1 (the pool pays its
    tables and scene bookkeeping), and the batched form's gain over the plain
    one comes from stepping many scenes per launch, not from a faster path."""
    from nonnewton_amd import autotune

    pool, slots = _admit_all([sc], 0)
    autotune.apply(pool, shapes)
    for _ in range(warmup):
        pool.tick()
    t = time.perf_counter()
    for _ in range(steps):
        pool.tick()
    pooled = time.perf_counter() - t
    pool.close()
    plain, _, _ = run_plain([sc], steps, warmup, shapes=shapes)
    return {"scene": sc.name, "dynamic_particles": sc.n_dynamic,
            "pooled_particle_steps_per_s": round(sc.n_dynamic * steps / pooled, 1),
            "plain_particle_steps_per_s": round(sc.n_dynamic * steps / plain, 1),
            "pooled_over_plain": round(plain / pooled, 3)}


def run_serving(count, rate, lanes=1, shapes=None):
    """Continuous batching over a Poisson request stream (sizes log-uniform 1k-500k,
    50-500 steps, seed 4), 200 GB budget, in `lanes` lanes."""
    from nonnewton_amd.serving import ServingLoop

    reqs = S.serving_requests(seed=4, rate=rate, count=count)
    loop = ServingLoop(200 << 30, lanes=lanes, shapes=shapes)
    rep = loop.run(reqs)
    loop.close()
    s = rep.summary()
    s["rate_per_s"] = rate
    s["lanes"] = lanes
    s["failed"] = int(sum(r.failed for r in rep.records))
    s["dynamic_particles_requested"] = int(sum(q.scene.n_dynamic for q in reqs))
    s["steps_requested"] = int(sum(q.steps for q in reqs))
    return s


def accuracy_checks(scs):
    """Against the oracle: (1) two scenes of the workload stepped 3 times on the
    GPU and in the FP32 oracle -- neighbor lists equal as sets and in order,
    fields bit for bit; (2) the dam break (13,824 fluid + 8,706 wall particles,
    dt 1e-3) 100 steps on the GPU against the FP64 oracle: the drift of the
    average density error (largest difference over the steps, % of rho0) and
    the largest error either run reached."""
    from oracle import oracle as O

    out = {}
    picks = sorted(range(len(scs)), key=lambda k: scs[k].n)[:2]
    same = True
    for k in picks:
        sc = scs[k]
        c = L.Context(sc.params, sc.n)
        c.upload(sc.x, sc.kind, sc.v)
        c.init()
        o = O.OracleSim.from_scene(sc, precision="f32")
        o.init()
        for _ in range(3):
            c.step()
            o.step()
        nb_g, c_g = c.neighbors()
        nb_o, c_o = o.neighbors()
        same &= bool(np.array_equal(c_g, c_o) and np.array_equal(nb_g, nb_o))
        for fg, fo in ((L.FIELD_X, O.F_X), (L.FIELD_V, O.F_V), (L.FIELD_DENSITY, O.F_DENSITY)):
            b = o.get(fo)
            a = c.download(fg).astype(np.float64).reshape(b.shape)
            if sc.meta.get("model", 0) == 0:
                same &= bool(np.array_equal(a, b))
            else:  # powf of the non-Newtonian models: last-bit differences between libraries
                same &= bool(np.abs(a - b).max() <= 1e-5 * max(np.abs(b).max(), 1e-12))
        c.close()
    out["neighbor_sets_and_fields_vs_fp32_oracle"] = {
        "scenes": [scs[k].name for k in picks], "steps": 3, "equal": same}
    db = S.dam_break()
    c = L.Context(db.params, db.n)
    c.upload(db.x, db.kind, db.v)
    c.init()
    o = O.OracleSim.from_scene(db, precision="f64")
    o.init()
    rho0 = db.params.density0
    drift = 0.0
    worst = 0.0
    for _ in range(100):
        sg = c.step()
        so = o.step()
        drift = max(drift, abs(sg["avg_density_err_p"] - so["avg_density_err_p"]))
        worst = max(worst, sg["avg_density_err_p"], so["avg_density_err_p"])
    c.close()
    out["dam_break_100_steps_vs_fp64"] = {
        "avg_density_error_drift_pct": float(f"{100 * drift / rho0:.3g}"),
        "max_avg_density_error_pct": float(f"{100 * worst / rho0:.3g}")}
    return out


def cpu_baseline(scs, budget_s, warm_steps=10):
    """The oracle in FP32 without compensated sums (the reference's arithmetic)
    on the host cores: `warm_steps` untimed steps per scene, then timed steps,
    round robin over the scenes, until about budget_s seconds of timed CPU
    work; the median scene-step time beside the aggregate rate."""
    from oracle import oracle as O

    cores = len(os.sched_getaffinity(0))
    O.set_threads(cores, "f32")
    done_dyn = 0
    spent = 0.0
    sims = []
    for s in scs:
        sim = O.OracleSim.from_scene(s, precision="f32", compensated=False)
        sim.init()
        for _ in range(warm_steps):
            sim.step()
        sims.append((s, sim))
    passes = 0
    per_scene = [[] for _ in sims]
    while spent < budget_s:
        # round robin over every scene, so the sample has the workload's mix
        for q, (s, sim) in enumerate(sims):
            t = time.perf_counter()
            sim.step()
            dt = time.perf_counter() - t
            spent += dt
            per_scene[q].append(dt)
            done_dyn += s.n_dynamic
        passes += 1
    steps = sum(len(v) for v in per_scene)
    med = [float(np.median(v)) for v in per_scene]
    return {
        "value": done_dyn / spent,
        "unit": "particle-steps/s",
        "cores": int(O.num_threads("f32")),
        "kind": "port",
        "median_scene_step_ms": round(1e3 * float(np.median(np.concatenate([np.asarray(v) for v in per_scene]))), 3),
        "median_round_ms": round(1e3 * sum(med), 1),
        "boundary_particles": int(sum(s.n_static for s, _ in sims)),
        "sample": f"oracle FP32 (reference arithmetic, OpenMP over particles): {warm_steps} untimed steps per scene, "
                  f"then {passes} round-robin pass(es) of one timed step over all {len(sims)} scenes of the workload; "
                  f"{steps} scene-steps, {done_dyn} particle-steps, {spent:.1f} s of timed CPU work; "
                  f"median_round_ms = sum over scenes of the median step time",
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


def roofline_entry(name, kernel_bytes, avg_s, launches, elapsed, workload):
    achieved = kernel_bytes / launches / avg_s / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(name, workload),
            "launches": launches, "avg_launch_us": round(1e6 * avg_s, 2),
            "bytes_per_launch": round(kernel_bytes / launches),
            "share_of_timed_region": round(avg_s * launches / elapsed, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=0, help="timed steps of the plain form (0 = --steps)")
    ap.add_argument("--workload", default="mixed_batch")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU leg (baseline and oracle checks)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-control", action="store_true", help="skip the one-scene pooled/plain control")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--lanes", type=int, default=3,
                    help="lanes of the batched form (pools stepped concurrently; 1 = one pool)")
    ap.add_argument("--trace-region", choices=("single", "lanes"), default="single",
                    help="which timed ticks roctxProfilerResume/Pause bracket: the one-lane form's (whose "
                         "launches carry the roofline's events) or the lanes'")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (no autotuning)")
    ap.add_argument("--serve-requests", type=int, default=64, help="requests per arrival rate (0 = no serving)")
    ap.add_argument("--serve-rates", default="5,20,80", help="Poisson arrival rates, requests/s")
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

    lanes = max(1, min(args.lanes, len(scs)))  # no lane without a scene
    b = run_batched(scs, args.steps, args.warmup, block=args.block, tune=not args.no_tune,
                    kernel=args.roofline_kernel,
                    region=TimedRegion() if (args.trace_region == "single" or lanes == 1) else _NoRegion())
    shapes = b["shapes"]
    # the single-lane form (one pool): its timed ticks carry the roofline's events
    s_elapsed, s_stats = b["elapsed"], b["stats"]
    s_value = n_dyn * args.steps / s_elapsed
    log(f"[bench] batched, one lane: {1e3 * s_elapsed / args.steps:.3f} ms/tick, {s_value:.4g} particle-steps/s")
    ln = None
    if lanes > 1:
        ln = run_lanes(scs, args.steps, args.warmup, lanes, block=args.block, shapes=shapes,
                       region=TimedRegion() if args.trace_region == "lanes" else None)
        elapsed, lat, stats = ln["elapsed"], ln["lat"], ln["stats"]
        log(f"[bench] batched, {lanes} lanes: {1e3 * elapsed / args.steps:.3f} ms per tick of every lane "
            f"(lanes {ln['lane_ms_per_tick']} ms/tick, scenes {ln['scenes_per_lane']})")
    else:
        elapsed, lat, stats = s_elapsed, b["lat"], s_stats
    value = n_dyn * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    log(f"[bench] batched: {ms_per_step:.3f} ms/tick, {value:.4g} particle-steps/s")
    if b["breakdown"]:
        log("[bench] kernel times of the probe tick (ms, launches): " +
            ", ".join(f"{k}={v[0]:.2f}/{v[1]}" for k, v in sorted(b["breakdown"].items(), key=lambda kv: -kv[1][0])
                      if v[1]))

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

    # roofline: the dominant kernel (the line's figure), a whole CG iteration
    # (matvec + update) and the whole tick, from algorithmic bytes over the
    # timed ticks and the event-timed mean launch durations
    roof = None
    extra = {}
    times = b["times"]
    if times:
        kb = {k: sum(kernel_algorithmic_bytes(k, sc, st) or 0 for ss in s_stats for sc, st in zip(scs, ss))
              for k in times}
        # a timed launch's interval holds the launch and event latency as well
        # as the kernel: measured on an empty kernel bracketed the same way
        ovh = L.event_overhead_ms(200) * 1e-3
        raw = {k: (ms * 1e-3 / timed) for k, (ms, timed, _) in times.items() if timed}
        avg = {k: max(v - ovh, 1e-9) for k, v in raw.items()}
        name = b["dominant"]
        if name in avg and kernel_algorithmic_bytes(name, scs[0], stats[0][0]) is not None:
            roof = roofline_entry(name, kb[name], avg[name], times[name][2], s_elapsed, args.workload)
            roof["timed_launches"] = times[name][1]
            roof["event_interval_us"] = round(1e6 * raw[name], 2)
            roof["event_overhead_us"] = round(1e6 * ovh, 2)
        if "cg_matvec" in avg and "cg_update" in avg:
            t_it = avg["cg_matvec"] * times["cg_matvec"][2] + avg["cg_update"] * times["cg_update"][2]
            by = kb["cg_matvec"] + kb["cg_update"]
            extra["cg_iteration"] = {"achieved": round(by / t_it / 1e9, 1),
                                     "frac": round(by / t_it / 1e9 / HBM_PEAK_GBS, 4),
                                     "bytes_per_tick": round(by / args.steps),
                                     "ms_per_tick": round(1e3 * t_it / args.steps, 3)}
            if name != "cg_update" and "cg_update" in avg:
                extra["cg_update"] = roofline_entry("cg_update", kb["cg_update"], avg["cg_update"],
                                                    times["cg_update"][2], s_elapsed, args.workload)
    # the whole tick: every sweep of every scene over the timed region, for the
    # line's form (lanes) and for the one-lane form the kernel figures come from
    for key, st_, el_ in (("tick", stats, elapsed), ("tick_single_lane", s_stats, s_elapsed)):
        tick_bytes = sum(step_algorithmic_bytes(sc, st) for ss in st_ for sc, st in zip(scs, ss))
        extra[key] = {"achieved": round(tick_bytes / el_ / 1e9, 1),
                      "frac": round(tick_bytes / el_ / 1e9 / HBM_PEAK_GBS, 4),
                      "bytes_per_tick": round(tick_bytes / args.steps)}
        if lanes == 1:
            break
    if roof is not None:
        roof.update(extra)
    log(f"[bench] roofline: {json.dumps(roof)}")

    it = np.array([[st["iterations_v"], st["iterations_p"], st["iterations_cg"]] for ss in stats for st in ss])
    rho0 = np.array([sc.params.density0 for sc in scs])
    err_p = np.array([[st["avg_density_err_p"] for st in ss] for ss in stats]) / rho0[None, :]
    nan = any(st["status"] & 8 for ss in stats for st in ss)
    cfg = {
        "workload": workload,
        "scenes": len(scs),
        "particles": n_all,
        "dynamic_particles": n_dyn,
        "form": f"{lanes} lanes" if lanes > 1 else "one pool",
        "batched_ms_per_tick": round(ms_per_step, 3),
        "p50_scene_step_ms_batched": round(1e3 * float(np.percentile(lat, 50)), 3),
        "p99_scene_step_ms_batched": round(1e3 * float(np.percentile(lat, 99)), 3),
        "mean_iterations_v_p_cg": [round(float(x), 2) for x in it.mean(axis=0)],
        "max_iterations_v_p_cg": [int(x) for x in it.max(axis=0)],
        # the pressure solve's average density error (FP64 reduction), worst
        # scene and tick of the timed region, in % of rho0
        "max_avg_density_error_pct": float(f"{100 * float(err_p.max()):.3g}"),
        "non_finite_state": bool(nan),
        "launch_shapes": shapes,
    }
    if ln is not None:
        cfg["lanes"] = {"lanes": lanes, "scenes": ln["scenes_per_lane"], "particles": ln["particles_per_lane"],
                        "ms_per_tick": ln["lane_ms_per_tick"]}
        cfg["one_lane_ms_per_tick"] = round(1e3 * s_elapsed / args.steps, 3)
        cfg["one_lane_particle_steps_per_s"] = round(s_value, 1)
    if plain_value is not None:
        cfg["plain_particle_steps_per_s"] = round(plain_value, 1)
        cfg["batched_over_plain"] = round(value / plain_value, 3)
        cfg["p50_scene_step_ms_plain"] = round(1e3 * float(np.percentile(plain_lat, 50)), 3)
        cfg["p99_scene_step_ms_plain"] = round(1e3 * float(np.percentile(plain_lat, 99)), 3)
        cfg["plain_steps_timed"] = plain_steps
        cfg["mean_iterations_v_p_cg_plain"] = [round(float(x), 2) for x in plain_its.mean(axis=0)]
    if not args.no_control and len(scs) > 1:
        mid = sorted(scs, key=lambda s: s.n_dynamic)[len(scs) // 2]
        cfg["one_scene_control"] = one_scene_control(mid, args.steps, args.warmup, shapes)
        log(f"[bench] one-scene control: {cfg['one_scene_control']}")

    if args.serve_requests > 0 and args.workload == "mixed_batch":
        cfg["serving"] = []
        for rate in [float(r) for r in args.serve_rates.split(",") if r]:
            s = run_serving(args.serve_requests, rate, lanes=lanes, shapes=shapes)
            cfg["serving"].append(s)
            log(f"[bench] serving at {rate:g}/s: {s}")

    cpu = None
    if not args.no_cpu:
        # the CPU leg: the oracle timed on the host cores (the baseline) and
        # used as the checker of the GPU results (accuracy)
        cpu = cpu_baseline(scs, args.cpu_seconds)
        log(f"[bench] cpu baseline: {cpu['value']:.4g} particle-steps/s on {cpu['cores']} cores")
        cfg["accuracy"] = accuracy_checks(scs)
        log(f"[bench] accuracy: {cfg['accuracy']}")

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
