"""The oracle's pinning tests catch plausible mistakes: each case below plants
one wrong coefficient, index or sign in a copy of oracle/sph_oracle.c, builds
it into a temporary directory and runs tests/test_oracle_forces.py against it
(ORACLE_LIB_DIR), which must then fail.  CPU only; the committed oracle is
not touched."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "sph_oracle.c")

# name: (text in the oracle, planted replacement)
MUTATIONS = {
    "weiler 2(d+2) = 8": ("real c = (real)10.0 * muij * (s->mass / rhoj)", "real c = (real)8.0 * muij * (s->mass / rhoj)"),
    "weiler without m_j/rho_j": ("real c = (real)10.0 * muij * (s->mass / rhoj)", "real c = (real)10.0 * muij"),
    "weiler mu_ij = mu_i": ("real muij = (real)0.5 * (mui + muj);", "real muij = mui;"),
    "weiler dt/rho0 for dt/rho_i": ("        real f = s->dt / s->rho[i];\n        for (int q = 0; q < 3; q++) {\n            real a = (acc",
                                    "        real f = s->dt / s->rho0;\n        for (int q = 0; q < 3; q++) {\n            real a = (acc"),
    "viscosity 2(d+2) = 8": ("const real d10 = (real)10.0 * s->nu;", "const real d10 = (real)8.0 * s->nu;"),
    "boundary viscosity m for rho0 V_j": ("real f = b10 * (s->rho0 * s->vol[j] / s->rho[i])", "real f = b10 * (s->mass / s->rho[i])"),
    "boundary reaction sign": ("s->a[3 * i + c] += -(ff * g[c]);", "s->a[3 * i + c] += (ff * g[c]);"),
    "elastic without the 1/2": ("fi[c] = (real)0.5 * fi[c];", "fi[c] = fi[c];"),
    "elastic dji sign": ("dji[c] = vi0 * g[c];", "dji[c] = -vi0 * g[c];"),
    "elastic V_i for V_j in f_ij": ("fij[c] = -vj0 * sdji[c];", "fij[c] = -vi0 * sdji[c];"),
}


@pytest.mark.parametrize("name", list(MUTATIONS))
def test_planted_mistake_is_caught(name, tmp_path):
    src = open(SRC).read()
    a, b = MUTATIONS[name]
    assert src.count(a) == 1, f"mutation site of {name!r} not found once in the oracle"
    mutated = tmp_path / "sph_oracle.c"
    mutated.write_text(src.replace(a, b))
    lib = tmp_path / "liboracle_f64.so"
    subprocess.check_call(["gcc", "-O1", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off",
                           "-DORACLE_F64", "-I", os.path.join(ROOT, "oracle"), str(mutated), "-o", str(lib), "-lm"])
    env = dict(os.environ, ORACLE_LIB_DIR=str(tmp_path), OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-x",
                        os.path.join(ROOT, "tests", "test_oracle_forces.py")], cwd=ROOT, env=env,
                       capture_output=True, text=True)
    assert r.returncode == 1 and "failed" in r.stdout, f"{name}: the pinning tests passed\n{r.stdout[-2000:]}"
