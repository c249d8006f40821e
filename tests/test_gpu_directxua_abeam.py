"""DirectXUA{0,0,1} with A-scaled beams on the device (AssemblyDirect with mh_set_A and
mh_beam_dRdA for the A column of the residual path) against the same solve with the oracle's
step assembly: every step's X and the identified A to 1e-9, and the device's ∂R/∂a against the
oracle's ln 10 · R of the masked section at a deformed state, 1e-12."""
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
