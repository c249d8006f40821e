"""DirectXUA{0,0,1} with A-scaled beams on the device (AssemblyDirect with mh_set_A and
mh_beam_dRdA for the A column of the residual path) against the same solve with the oracle's
step assembly: every step's X and the identified A to 1e-9, and the device's ∂R/∂a against the
oracle's ln 10 · R of the masked section at a deformed state, 1e-12. With every step in one launch
(batch_steps: the A-dofs shared by the steps' copies) the iterates are bitwise those of one
launch per step, and with the all-steps system and its LU on the device they agree to 1e-9."""
import numpy as np
import pytest

from tests.helpers import OracleAssemblyDirect
from tests.test_directxua_abeam import A_TRUE, abeam_model, identify, truth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def test_device_dRdA_matches_oracle(mc):
    model = abeam_model()
    st = mc.initialize(model, time=0.)
    rng = np.random.default_rng(5)
    st.X[0][:] = rng.normal(0., 1e-2, st.X[0].size)
    st.A[:] = [0.1, -0.2]
    dev = mc.AssemblyDirect(0, 0, 1, model, st.dis, skip=(1, 2, 3, 4, 5, 6, 7))
    dev.assemble(st, np.zeros(model.ndof["X"]))
    D = dev.beam_dRdA(0)
    oa = OracleAssemblyDirect(0, 0, 1, model, st.dis)
    oa._state = st
    Do = oa.beam_dRdA(0)
    assert np.abs(D - Do).max() <= 1e-12 * np.abs(Do).max()


def test_directxua_abeam_device_matches_oracle(mc):
    meas = truth()
    Ad, outd = identify(meas, Assembly=mc.AssemblyDirect)
    Ao, outo = identify(meas, Assembly=OracleAssemblyDirect)
    assert np.abs(Ad - Ao).max() <= 1e-9 * np.abs(Ao).max()
    assert np.abs(Ad - A_TRUE).max() <= 1e-3 * np.abs(A_TRUE).max()
    for sd, so in zip(outd[0], outo[0]):
        assert np.abs(sd.X[0] - so.X[0]).max() <= 1e-9 * np.abs(so.X[0]).max()


def test_abeam_steps_in_one_launch(mc):
    meas = truth()
    A1, o1 = identify(meas, Assembly=mc.AssemblyDirect)
    A2, o2 = identify(meas, Assembly=mc.AssemblyDirect, solve_kw=dict(batch_steps=True))
    assert np.array_equal(A1, A2)
    for s1, s2 in zip(o1[0], o2[0]):
        assert np.array_equal(s1.X[0], s2.X[0])
    A3, o3 = identify(meas, Assembly=mc.AssemblyDirect, solve_kw=dict(batch_steps=True, big="device",
                                                                      linsolve="device"))
    assert np.abs(A3 - A1).max() <= 1e-9 * np.abs(A1).max()
    for s1, s3 in zip(o1[0], o3[0]):
        assert np.abs(s3.X[0] - s1.X[0]).max() <= 1e-9 * np.abs(s1.X[0]).max()
