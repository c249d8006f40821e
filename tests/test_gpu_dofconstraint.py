"""DofConstraint{:X} with user gap functions in solve(SweepX{0}) on the device (host-evaluated
blocks, mh_set_host_blocks, scattered with the device kinds): the interior-point case of
TestDofConstraints.jl:162-182 (two positive-mode constraints with nonlinear gaps, gravity):
both constraints end active, |g1| < 1e-6 and |g2| < 1e-5; the same with the device linear
solve (host blocks read the device-resident X)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def g1(x, t):
    return -0.1 * np.sin(1.2 * x[0]) + .2 * x[0] + x[1] + 1


def g2(x, t):
    return -.4 * x[0] + .3 * x[1] + .1


def model_of(mc):
    model = mc.Model("interior")
    n1 = mc.addnode(model, [0., 0.])
    mc.addelement(model, mc.DofConstraint, [n1], xinod=(1, 1), xfield=("t1", "t2"), λinod=1, λclass="X",
                  λfield="λ1", gap=g1, mode=mc.positive)
    mc.addelement(model, mc.DofConstraint, [n1], xinod=(1, 1), xfield=("t1", "t2"), λinod=1, λclass="X",
                  λfield="λ2", gap=g2, mode=mc.positive)
    mc.addelement(model, mc.DofLoad, [n1], field="t2", value=lambda t: -2.)
    st = mc.initialize(model)
    st = mc.setdof(st, 1., field="λ1")
    st = mc.setdof(st, 1., field="λ2")
    return model, st


def test_interior_point(mc):   # TestDofConstraints.jl:162-182
    model, st = model_of(mc)
    out = mc.solve(mc.SweepX[0], initialstate=st, time=[0.], verbose=False, catcherror=False)
    X = out[-1].X[0][:2]
    assert abs(g1(X, 0)) < 1e-6 and abs(g2(X, 0)) < 1e-5
    lam = out[-1].X[0][2:]
    assert np.all(lam > 0)


def test_interior_point_device_linsolve(mc):
    model, st = model_of(mc)
    out = mc.solve(mc.SweepX[0], initialstate=st, time=[0.], verbose=False, catcherror=False, linsolve="device")
    X = out[-1].X[0][:2]
    assert abs(g1(X, 0)) < 1e-6 and abs(g2(X, 0)) < 1e-5
