"""Host side of element results (muscade_amd/results.py, StrainGaugeOnEulerBeam3D): the gauge
constructor's E, K1, K2, K3 against TestBeamElementStrainGauge.jl:13-27, the P[1,:] = 0 rule,
and the layout split of mh_beam_results (the oracle's identical layout)."""
import math

import numpy as np
import pytest

import muscade_amd as mc
from muscade_amd.results import _split
from tests import _oracle as orc
from tests.test_oracle_beam import mat16

P = np.array([[0., .5, 0.], [0., 0, .5], [0., -.5, 0.], [0., 0, -.5], [0., .5, 0.]]).T
D = np.array([[1., 0., 0.], [1., 0., 0.], [1., 0., 0.], [1., 0., 0.], [1 / math.sqrt(2), 0, 1 / math.sqrt(2)]]).T


def gauge_model(Pm=P):
    model = mc.Model("gauge")
    n1 = mc.addnode(model, [0., 0., 0.])
    n2 = mc.addnode(model, [4., 0., 0.])
    mat = mc.BeamCrossSection(EA=10., EI2=3., EI3=3., GJ=4., mu=1., iota1=1.0)
    mc.addelement(model, mc.StrainGaugeOnEulerBeam3D, [n1, n2], P=Pm, D=D,
                  elementkwargs=dict(mat=mat, orient2=np.array([0., 1., 0.])))
    return model


def test_gauge_constructor_goldens():
    model = gauge_model()
    g = model.eletypes[0].params[0]["gauges"][0]
    assert np.allclose(g[:, 0], [1., 1., 1., 1., .5]) and np.allclose(g[:, 1], [0., 0., 0., 0., .25])
    assert np.allclose(g[:, 2], [-0.5, 0, 0.5, 0, -.25]) and np.allclose(g[:, 3], [0, -0.5, 0, 0.5, 0])
    assert model.ndof["X"] == 12
    bad = P.copy()
    bad[0, 1] = 0.1
    with pytest.raises(mc.MuscadeError):
        gauge_model(bad)


@pytest.mark.parametrize("ND", [1, 3])
def test_split_layout(ND):
    rng = np.random.default_rng(4)
    X = [rng.normal(size=12) * 0.2 for _ in range(ND)]
    m = mat16(7., 3., 2., 4., 1.5, 0.7, w=2.5)
    raw = orc.beam_results([0, 0, 0], [3, 1, 0], [0, 0, 1.], m, X)
    r = _split(raw[None, :], ND, 0)
    assert r["ε"].shape == (1, ND) and r["κ"].shape == (1, ND, 3)
    assert r["gp"]["x"].shape == (1, 4, ND, 3) and r["gp"]["resultants"]["mₑ"].shape == (1, 4, 3)
    # fᵢ = EA·ε and mᵢ = (GJ, EI₃, EI₂)·κgp at every Gauss point
    assert np.allclose(r["gp"]["resultants"]["fᵢ"][0], 7. * r["ε"][0, 0], rtol=1e-15)
    k = r["gp"]["κgp"][0, :, 0]
    assert np.allclose(r["gp"]["resultants"]["mᵢ"][0], k * np.array([4., 2., 3.]), rtol=1e-15)
    sel = mc.results._select(r, mc.request("κ", gp=mc.request(resultants=mc.request("fₑ"))), "EulerBeam3D")
    assert set(sel) == {"κ", "gp"} and set(sel["gp"]) == {"resultants"} and set(sel["gp"]["resultants"]) == {"fₑ"}
