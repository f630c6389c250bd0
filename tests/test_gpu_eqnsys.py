"""The library's equation systems (dfm_eqn_*: CEqnSystem_Solid2D, CEqn_Solid3D_Linear, CEqnSystem_Scalar2D,
CEqn_Scalar3D) against the restated Solve flows (oracle EqnFlow, itself pinned bit for bit to the reference's own
classes by tests/test_oracle_eqnsys.py), scenario by scenario: the same number of PCG solves, each within 2
iterations of the oracle's and below conv_ratio 1e-6, and the field values close to the oracle's. The device PCG
sums in another order than the oracle's, so its iterates differ by rounding: where every solve takes the
oracle's iteration count the fields agree to 1e-8 relative, else to the solver tolerance times the condition
number (1e-4 here)."""
import numpy as np
import pytest

from eqn_cases import SCENARIOS
from oracle import pyref as P

pytestmark = pytest.mark.gpu

N = None


@pytest.fixture(scope="module")
def ctx():
    global N
    from delfem_amd import _native

    N = _native
    c = _native.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_eqn_matches_oracle_flow(ctx, name):
    print(f"equation system {name}", flush=True)
    gpu = SCENARIOS[name](lambda *a: N.Eqn(ctx, *a))
    cpu = SCENARIOS[name](P.EqnFlow)
    gi, ci = gpu.itr(), cpu.itr()
    assert len(gi) == len(ci), (gi, ci)
    same = True
    for (g_it, g_c), (c_it, c_c) in zip(gi, ci):
        assert abs(g_it - c_it) <= 2, (gi, ci)
        assert g_c < 1e-6 and c_c < 1e-6
        same = same and g_it == c_it
    tol = 1e-8 if same else 1e-4
    for k, (a, b) in enumerate(zip(gpu.fields(), cpu.fields())):
        nb = np.linalg.norm(b)
        if nb == 0.0:
            assert np.linalg.norm(a) == 0.0, k
            continue
        rel = np.linalg.norm(a - b) / nb
        print(f"  field {k}: rel diff {rel:.2e} (iterations {gi} vs {ci})")
        assert rel <= tol, (k, rel)
    gpu.close()


def test_flags_and_errors(ctx):
    from eqn_cases import mesh2d, mesh3d

    co, tri, group, bnd = mesh2d()
    e = N.Eqn(ctx, N.EQN_SOLID2D, co, tri, group, 2, bnd, -1)
    with pytest.raises(N.DfmError):
        e.set_scalar(0, 1.0)  # not a scalar equation
    with pytest.raises(N.DfmError):
        e.set_young_poisson(1.0, 0.3, True, 5)  # no such group
    e.set_flag(1, False)
    e.set_flag(2, True)  # a dynamic system with the saved matrix: not built
    with pytest.raises(N.DfmError):
        e.solve()
    e.close()
    co3, tet, _, b3 = mesh3d()
    with pytest.raises(N.DfmError):
        N.Eqn(ctx, N.EQN_SOLID3D, co3, tet, np.zeros(tet.size // 4, np.int32), 2, b3, -1)
