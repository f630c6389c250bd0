"""The restated Solve flows of the equation systems (oracle/pyref.py EqnFlow) against the reference's own
CEqnSystem_Solid2D / CEqn_Solid3D_Linear / CEqnSystem_Scalar2D / CEqn_Scalar3D (oracle/_ref through
oracle/ref_capi.cpp), bit for bit: the field values after each scenario and GetAry_ItrNormRes."""
import numpy as np
import pytest

from eqn_cases import SCENARIOS
from oracle import pyref as P

pytestmark = pytest.mark.skipif(not P.ref_available(), reason="oracle/_ref not built")


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_flow_matches_reference(name):
    ref = SCENARIOS[name](P.RefEqn)
    flow = SCENARIOS[name](P.EqnFlow)
    assert flow.itr() == ref.itr(), (flow.itr(), ref.itr())
    assert len(ref.itr()) >= 1 and all(c < 1e-6 for _, c in ref.itr())
    for k, (a, b) in enumerate(zip(flow.fields(), ref.fields())):
        if k > 0 and not np.any(b) and not np.any(a):
            continue
        assert np.array_equal(a, b), (k, np.abs(a - b).max())
    u = ref.fields()[0]
    assert np.abs(u).max() > 0.0


def test_plane_stress_flag_is_inverted():
    """SetYoungPoisson(..., is_plane_stress = true) keeps the 3D lambda (plane strain); false reduces it."""
    lam3 = 1.0 * 0.3 / ((1.0 + 0.3) * (1 - 2.0 * 0.3))
    myu = 1.0 / (2.0 * 1.3)
    assert P._solid_lame(1.0, 0.3, True) == (lam3, myu)
    lam_s, _ = P._solid_lame(1.0, 0.3, False)
    assert np.isclose(lam_s, 1.0 * 0.3 / (1 - 0.3 ** 2))  # E nu / (1 - nu^2): plane stress
