"""The all-steps DirectXUA system factorized on the device (DeviceBigSystem.solve:
mh_big_solve_prepare / mh_big_solve, in place of lu(Lvv) and Lvv\\Lv, src/DirectXUA.jl:489-490)
against SciPy's sparse LU of the same matrix (downloaded from the same device assembly):
- each iteration's Δv: normwise backward error ≤ 1e-13 and at most SuperLU's (×10; the
  systems are too ill-conditioned for a forward-error comparison);
- solve(DirectXUA{OX,0,IA}; big="device", linsolve="device") reaches the linsolve="host"
  history (within one iteration, states to 1e-8), for OX = 0 (block diagonal), OX = 1 and 2 (two
  steps per diagonal block), and IA = 1 (the A-dof borders every step's block);
- a singular system is reported as DirectXUA's factorization failure, and blocks too large for
  the dense solve are refused at prepare."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from tests.test_gpu_directxua import beam_model

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def with_A(mc, model, nel=6):
    """An A-dof a scaling a target tip motion, the same for every step (IA = 1)."""
    mc.addelement(model, mc.DofCost, [nel + 1], xinod=(1,), xfield=("t2",), ainod=(1,), afield=("a",),
                  cost=lambda X, U, A, t: 1e2 * (X[0][0] - 1e-3 * A[0] * np.sin(t)) ** 2)
    # the prior enters once (assembleA!) and at every step (the reference's element loop reaches
    # Acost elements too), so its weight is kept small
    mc.addelement(model, mc.SingleAcost, [nel + 1], field="a", cost=lambda a: 1e-3 * (a - 1.) ** 2)
    return model


def run_pair(mc, model, OX, IA, time, spy=None):
    st0 = mc.initialize(model, time=0.)
    kw = dict(initialstate=[st0], time=[time], maxΔx=1e-9, maxΔλ=1e-6, maxΔu=1e-6, maxΔa=1e-6, maxiter=30,
              verbose=False, catcherror=False, batch_steps=True, big="device")
    sh, sd = {}, {}
    host = mc.solve(mc.DirectXUA[OX, 0, IA], linsolve="host", stats=sh, **kw)
    dev = mc.solve(mc.DirectXUA[OX, 0, IA], linsolve="device", stats=sd, **kw)
    return host, dev, sh, sd


@pytest.mark.parametrize("OX,inverse,IA", [(0, True, 0), (0, True, 1), (1, False, 0), (2, False, 0), (2, True, 0),
                                           (2, True, 1)])
def test_device_solve_reaches_host_history(mc, OX, inverse, IA):
    print("case", OX, inverse, IA, flush=True)
    model = beam_model(mc, dyn_load=not inverse, inverse=inverse)
    if IA:
        model = with_A(mc, model)
    time = np.arange(1., 10.) * (0.1 if OX else 1.)
    host, dev, sh, sd = run_pair(mc, model, OX, IA, time)
    dX = max(np.abs(a.X[d] - b.X[d]).max() / (np.abs(b.X[d]).max() + 1e-300)
             for a, b in zip(dev[0], host[0]) for d in range(OX + 1))
    dU = max(abs(a.U[0][0] - b.U[0][0]) / abs(b.U[0][0]) for a, b in zip(dev[0], host[0])) if inverse else 0.
    print("iterations", sh["iterations"], sd["iterations"], "rel. diff X", dX, "U", dU, flush=True)
    # the two factorizations round differently: the last Newton step may fall on either side of
    # the tolerance, one iteration apart
    assert abs(sh["iterations"] - sd["iterations"]) <= 1
    if OX:
        assert sd["blocks"] == 5 and sd["block_order"] == 2 * (2 * model.ndof["X"] + model.ndof["U"])
    assert dX <= 1e-8 and dU <= 1e-6
    if IA:
        ad, ah = dev[0][0].A[0], host[0][0].A[0]
        print("A", ad, ah, flush=True)
        assert abs(ad - ah) <= 1e-8 * abs(ah) and abs(ah - 1.) > 1e-3


def test_each_solve_against_splu(mc):
    from muscade_amd import directxua as dx
    model = with_A(mc, beam_model(mc, inverse=True))
    time = np.arange(1., 8.) * 0.1
    errs = []
    orig = dx.DeviceBigSystem.solve

    def spy(self, hcontrib, rhs, dbg=None):
        x = orig(self, hcontrib, rhs, dbg)
        M = self.assemble(hcontrib)
        def berr(y):
            return np.abs(M @ y - rhs).max() / (abs(M).max() * np.abs(y).sum() + np.abs(rhs).max())
        xs = spla.spsolve(M.tocsc(), rhs)
        errs.append((berr(x), berr(xs), np.abs(x - xs).max() / np.abs(xs).max()))
        return x
    dx.DeviceBigSystem.solve = spy
    try:
        st0 = mc.initialize(model, time=0.)
        mc.solve(mc.DirectXUA[2, 0, 1], initialstate=[st0], time=[time], maxΔx=1e-9, maxΔλ=1e-6, maxΔu=1e-6,
                 maxΔa=1e-6, maxiter=30, verbose=False, catcherror=False, batch_steps=True, big="device",
                 linsolve="device")
    finally:
        dx.DeviceBigSystem.solve = orig
    assert errs
    for be, bs, re in errs:
        print("backward error device", be, "SuperLU", bs, "relative difference", re, flush=True)
        # the all-steps systems of this inverse problem are ill-conditioned (the two solutions
        # differ by up to 1e-3 relative while both are backward stable): compare backward errors
        assert be <= 1e-13 and be <= 10. * bs + 1e-20, (be, bs)


@pytest.mark.parametrize("nel", [48, 123])
def test_large_blocks_against_splu(mc, nel):
    """Block orders past one panel row slot of the device LU: nel = 48 gives m = 1,200 (rows
    1024..2400 of a block fall in the panel kernel's register slots 1 and 2), nel = 123 gives
    m = 3,000 (above 2,560: the rocSOLVER getrf path). Backward error of every Newton system."""
    from muscade_amd import directxua as dx
    model = beam_model(mc, nel=nel, dyn_load=True)
    time = np.arange(1., 8.) * 0.1
    errs, orders = [], []
    orig = dx.DeviceBigSystem.solve

    def spy(self, hcontrib, rhs, dbg=None):
        x = orig(self, hcontrib, rhs, dbg)
        M = self.assemble(hcontrib)

        def berr(y):
            return np.abs(M @ y - rhs).max() / (abs(M).max() * np.abs(y).sum() + np.abs(rhs).max())
        xs = spla.spsolve(M.tocsc(), rhs)
        errs.append((berr(x), berr(xs)))
        return x
    dx.DeviceBigSystem.solve = spy
    try:
        st0 = mc.initialize(model, time=0.)
        sd = {}
        mc.solve(mc.DirectXUA[2, 0, 0], initialstate=[st0], time=[time], maxΔx=1e-9, maxΔλ=1e-6, maxiter=6,
                 verbose=False, catcherror=True, silenterror=True, batch_steps=True, big="device", linsolve="device",
                 stats=sd)
    finally:
        dx.DeviceBigSystem.solve = orig
    print("nel", nel, "block order", sd.get("block_order"), "solves", len(errs), flush=True)
    assert sd["block_order"] == 2 * (2 * model.ndof["X"] + model.ndof["U"])
    assert errs
    for be, bs in errs:
        print("backward error device", be, "SuperLU", bs, flush=True)
        assert be <= 1e-13 and be <= 10. * bs + 1e-20, (be, bs)


def test_singular_system_raises(mc):
    """An X-dof that only a zero cost touches: its column (and its Λ row) is exactly zero."""
    model = beam_model(mc, dyn_load=True)
    n = mc.addnode(model, [10., 0., 0.])
    mc.addelement(model, mc.DofCost, [n], xinod=(1,), xfield=("t1",), cost=lambda X, U, A, t: 0. * X[0][0] * X[0][0])
    st0 = mc.initialize(model, time=0.)
    with pytest.raises(mc.MuscadeError, match="factorization failed"):
        mc.solve(mc.DirectXUA[0, 0, 0], initialstate=[st0], time=[np.arange(1., 4.)], maxiter=3, verbose=False,
                 catcherror=False, batch_steps=True, big="device", linsolve="device")


def test_too_large_blocks_are_refused(mc):
    """Diagonal blocks beyond the dense solve's limit (m > 4000) are refused at prepare with
    MH_ERR_SIZE, before any factorization."""
    model = beam_model(mc, nel=400, dyn_load=True)
    st0 = mc.initialize(model, time=0.)
    with pytest.raises(mc.MuscadeError, match="too large"):
        mc.solve(mc.DirectXUA[2, 0, 0], initialstate=[st0], time=[np.arange(1., 8.) * 0.1], maxiter=2,
                 verbose=False, catcherror=False, batch_steps=True, big="device", linsolve="device")
