"""The FP32-flux precision form against the plain FP64 form on the GPU.

The FP32-flux form evaluates the MUSCL reconstruction, the limiters and the
face fluxes of tilde (momentum), ygas (species) and htilde (enthalpy) in
single precision and stores those face fluxes in single precision; the
pressure-velocity iteration (its mass fluxes included), the temperature
inversion and every accumulation stay FP64.  The bar of the form: relative L2
error of p and of each particle volume fraction
(rlk / rho_s) at most 1e-4 against the FP64 form after the same steps; the
error of every state field is checked against a looser 1e-3 and reported.
"""
import numpy as np
import pytest

from pdac_amd import abi
from tests.golden.make_golden import CASES, load

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a - b)


def run(case, precision, nsteps):
    from pdac_amd.gpu import gpu_solver
    s = gpu_solver(case, device=0, precision=precision)
    s.step(nsteps)
    out = s.state()
    s.close()
    return out


def errors(case, a, b):
    e = {k: rel_l2(a[k], b[k]) for k in b}
    rho = np.asarray(case.density[:case.nsolid]).reshape(-1, 1)
    e["eps_s"] = max(rel_l2(a["rlk"].reshape(case.nsolid, -1)[s] / rho[s],
                            b["rlk"].reshape(case.nsolid, -1)[s] / rho[s])
                     for s in range(case.nsolid))
    return e


@pytest.mark.parametrize("name", [c for c in sorted(CASES)
                                  if c != "mixed_2d_limiters" and not c.startswith("explicit_")])
def test_fp32flux_fixture_error(name):
    case, _, _ = load(name)
    a = run(case, abi.PREC_FP32FLUX, case.nsteps)
    b = run(case, abi.PREC_FP64, case.nsteps)
    e = errors(case, a, b)
    print(name, " ".join(f"{k}:{v:.1e}" for k, v in sorted(e.items())), flush=True)
    assert e["p"] <= 1e-4 and e["eps_s"] <= 1e-4, e
    assert max(e.values()) <= 1e-3, e
    assert max(e.values()) > 0.0  # the forms differ: the FP32 kernels ran


def test_fp32flux_eruption_column_200_steps():
    from pdac_amd.cases import eruption_column
    case = eruption_column(nx=128, nz=256, nsteps=200)
    a = run(case, abi.PREC_FP32FLUX, 200)
    b = run(case, abi.PREC_FP64, 200)
    e = errors(case, a, b)
    print("eruption 128x256, 200 steps:", " ".join(f"{k}:{v:.1e}" for k, v in sorted(e.items())),
          flush=True)
    assert e["p"] <= 1e-4 and e["eps_s"] <= 1e-4, e
