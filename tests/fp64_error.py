"""Error of the FP32 HIP step against the FP64 oracle on the same inputs.

For the dam break (13,824 fluid + 8,706 wall particles, dt 1e-3), the
non-Newtonian dam break (131,072 fluid, Cross model, Weiler CG, dt 5e-4) and
the elastic drop (98,304 elastic + 40,000 plate, dt 2e-4):

- single step, per field: max |a - b|, RMS(a - b) and max |b| (b the FP64
  oracle), for one step from the initial state and for one step from the GPU's
  state late in a run (the oracle, initialized on the initial positions,
  continues from the GPU's x, v and Weiler warm-start direction);
- drift of the average density error over 100 steps: the largest difference
  between the GPU's and the FP64 oracle's per-step average density error of the
  pressure solve (SPH/main.py:1669-1675), in % of rho0, for the dam break and
  for a fluid block thrown at the floor (its pressure solve iterates every
  step; the dam breaks stay below rest density).

    python tests/fp64_error.py [--out table.md]      (needs a GPU)

tests/test_gpu_accuracy.py asserts bounds on the same figures.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

from nonnewton_amd import _lib as L  # noqa: E402
from nonnewton_amd import scenes  # noqa: E402
from oracle import oracle as O  # noqa: E402

FLUID = [("x", L.FIELD_X, O.F_X), ("v", L.FIELD_V, O.F_V), ("acceleration", L.FIELD_ACCELERATION, O.F_ACC),
         ("density", L.FIELD_DENSITY, O.F_DENSITY), ("dfsph_factor", L.FIELD_DFSPH_FACTOR, O.F_DFSPH_FACTOR)]
NONNEWTON = [("viscosity", L.FIELD_VISCOSITY, O.F_VISCOSITY), ("strain_norm", L.FIELD_STRAIN_NORM, O.F_STRAIN_NORM)]
ELASTIC = [("x", L.FIELD_X, O.F_X), ("v", L.FIELD_V, O.F_V), ("acceleration", L.FIELD_ACCELERATION, O.F_ACC),
           ("rotation", L.FIELD_ROTATION, O.F_ROTATION), ("stress", L.FIELD_STRESS, O.F_STRESS)]

WORKLOADS = {
    "dam_break": (scenes.dam_break, FLUID, 50),
    "nonnewton_dam": (scenes.nonnewton_dam, FLUID + NONNEWTON, 50),
    "elastic_drop": (scenes.elastic_drop, ELASTIC, 100),
}


def field_errors(ctx, orc, fields, mask=None):
    out = {}
    for name, fg, fo in fields:
        b = orc.get(fo)
        a = ctx.download(fg).astype(np.float64).reshape(b.shape)
        if mask is not None:
            a, b = a[mask], b[mask]
        d = np.abs(a - b)
        out[name] = {"max": float(d.max()), "rms": float(np.sqrt(np.mean(d * d))), "scale": float(np.abs(b).max())}
    return out


def single_step(name, late=True):
    """{'initial': errors, 'late': errors} for workload `name`."""
    make, fields, k_late = WORKLOADS[name]
    sc = make()
    mask = sc.kind != 1 if name != "elastic_drop" else sc.kind == 2  # dynamic particles
    res = {}
    ctx = L.Context(sc.params, sc.n)
    ctx.upload(sc.x, sc.kind, sc.v)
    ctx.init()
    orc = O.OracleSim.from_scene(sc, precision="f64")
    orc.init()
    ctx.step()
    orc.step()
    res["initial"] = field_errors(ctx, orc, fields, mask)
    del orc
    if late:
        for _ in range(k_late - 1):
            ctx.step()
        orc = O.OracleSim.from_scene(sc, precision="f64")
        orc.init()
        for f, fo in ((L.FIELD_X, O.F_X), (L.FIELD_V, O.F_V), (L.FIELD_VISC_DIFF, O.F_VISC_DIFF)):
            orc.set(fo, ctx.download(f))
        ctx.step()
        orc.step()
        res[f"after step {k_late}"] = field_errors(ctx, orc, fields, mask)
    ctx.close()
    return res


def drift(sc, steps=100):
    """Largest |err_gpu - err_fp64| of the pressure solve's average density
    error over `steps` steps, and the largest error either reached (% of rho0)."""
    ctx = L.Context(sc.params, sc.n)
    ctx.upload(sc.x, sc.kind, sc.v)
    ctx.init()
    orc = O.OracleSim.from_scene(sc, precision="f64")
    orc.init()
    rho0 = sc.params.density0
    d = worst = 0.0
    for _ in range(steps):
        sg = ctx.step()
        so = orc.step()
        d = max(d, abs(sg["avg_density_err_p"] - so["avg_density_err_p"]))
        worst = max(worst, sg["avg_density_err_p"], so["avg_density_err_p"])
    ctx.close()
    return {"drift_pct": 100 * d / rho0, "max_err_pct": 100 * worst / rho0}


def table(out=None):
    lines = ["| workload | state | field | max abs error | RMS error | max abs value (FP64) | max / value |",
             "|---|---|---|---|---|---|---|"]
    allres = {}
    for name in WORKLOADS:
        r = single_step(name)
        allres[name] = r
        for state, errs in r.items():
            for f, e in errs.items():
                rel = e["max"] / e["scale"] if e["scale"] > 0 else 0.0
                lines.append(f"| {name} | {state} | {f} | {e['max']:.3e} | {e['rms']:.3e} | {e['scale']:.3e} | {rel:.2e} |")
    dr = {"dam_break": drift(scenes.dam_break()), "impact_block": drift(scenes.impact_block())}
    allres["drift"] = dr
    lines.append("")
    lines.append("| workload | 100-step drift of the average density error (% of rho0) | largest average density error (% of rho0) |")
    lines.append("|---|---|---|")
    for k, v in dr.items():
        lines.append(f"| {k} | {v['drift_pct']:.3e} | {v['max_err_pct']:.3e} |")
    text = "\n".join(lines) + "\n"
    if out:
        with open(out, "w") as f:
            f.write(text)
    return allres, text


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    _, t = table(a.out)
    print(t, flush=True)
