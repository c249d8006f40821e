"""Edge cases of the device assembly against the oracle's sequential assembler (bit-exact
pattern, values within 1e-12 normwise):
- a single element of each kind and the element counts around a workgroup's size (the static
  beam kernel's 128, the dynamic direction pass's 9 elements per workgroup, the bar kernel's
  128), so that the last workgroup is ragged or exactly full;
- a model of point elements only (Hold and DofLoad, no beam or bar kernel launched)."""
import numpy as np
import pytest

from tests.helpers import oracle_model, relerr
from tests.test_gpu_assembly import random_beams
from tests.test_gpu_bar import random_bars
from tests.test_gpu_dynamic import DT, dyn_beams

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def check(mc, model, state, OX, mission, udof=False):
    out = mc.AssemblySweepX(OX, model, state.dis)
    if OX:
        out.c = mc.Newmarkβcoefficients(OX, DT, 0.25, 0.5)
        c = out.c.array()
    else:
        c = np.zeros(6)
    Lλ, nzval = out.assemble(state, mission)
    om = oracle_model(model, state.dis, t=state.time)
    r, rhs, nz, err = om.assemble(OX, 0 if mission == "step" else 1, c, state.X[:OX + 1],
                                  state.U[0] if udof else None, None)
    assert r == 0, err
    assert np.array_equal(out.colptr, om.colptr) and np.array_equal(out.rowval, om.rowval)
    assert relerr(Lλ, rhs) <= TOL
    assert relerr(nzval, nz) <= TOL


@pytest.mark.parametrize("n", [1, 2, 127, 128, 129, 256])
def test_static_beams_ragged(mc, n):
    print("static beams", n, flush=True)
    model, state, *_ = random_beams(mc, n, seed=100 + n)
    check(mc, model, state, 0, "iter")


@pytest.mark.parametrize("n", [1, 8, 9, 10, 18])
def test_dynamic_beams_ragged(mc, n):
    print("dynamic beams", n, flush=True)
    model, state, *_ = dyn_beams(mc, n, seed=200 + n, rotscale=0.4)
    check(mc, model, state, 2, "step")


@pytest.mark.parametrize("n", [1, 128, 129])
def test_bars_ragged(mc, n):
    print("bars", n, flush=True)
    model, state, *_ = random_bars(mc, n, seed=300 + n)
    check(mc, model, state, 2, "iter")


def test_point_elements_only(mc):
    model = mc.Model("points")
    mc.addnode(model, [[0., 0., 0.], [1., 0., 0.]])
    for f in ("t1", "t2", "t3"):
        mc.addelement(model, mc.Hold, [1], field=f)
    mc.addelement(model, mc.DofLoad, [2], field="t1", value=lambda t: 3. * t)
    state = mc.initialize(model, time=2.)
    rng = np.random.default_rng(4)
    state.X[0][:] = rng.normal(size=model.ndof["X"])
    check(mc, model, state, 0, "iter")
