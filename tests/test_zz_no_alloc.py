"""mh_beam_gauge_derivs_noalloc (the gauge derivatives with no allocation, out of a fixed
device scratch array) against mh_beam_gauge_derivs, through the C ABI: bitwise equal on two
calls in a row, the second with a smaller selection, on all elements and on a selection;
sizes beyond the scratch are refused on the host with MH_ERR_SIZE."""
import numpy as np
import pytest

from tests.helpers import build_lattice

pytestmark = pytest.mark.gpu

GAUGE_OUT = 364          # values, gradients and Hessians of εₐₓ, κ₁, κ₂, κ₃ per element
NOALLOC_MAX = 4096       # elements per call of the no-allocation entry point


def _beam_context(nx, ny, nz, seed):
    import muscade_amd as mc
    from muscade_amd import _lib
    model, state = build_lattice(nx, ny, nz, seed=seed, rot=2e-2, disp=1e-3)
    out = mc.AssemblySweepX(0, model, state.dis)
    beams = [it for it, E in enumerate(model.eletypes) if E.cls.kind == _lib.MH_KIND_BEAM]
    assert beams
    out.ctx.check(_lib.lib.mh_set_X(out.ctx.h, 0, _lib.dptr(_lib.as_f64(state.X[0]))), "set X")
    return out, model, beams[0]


def _call(out, name, group, sel, n):
    from muscade_amd import _lib
    res = np.full((n, GAUGE_OUT), np.nan)
    s = None if sel is None else _lib.as_i64(sel)
    r = getattr(_lib.lib, name)(out.ctx.h, out.groups[group], 0 if s is None else s.size,
                                _lib.i64ptr(s), _lib.dptr(res))
    out.ctx.check(r, name)
    return res


def test_noalloc_equals_allocating_entry_point():
    out, model, it = _beam_context(4, 4, 3, seed=11)
    nele = model.eletypes[it].nele
    assert nele <= NOALLOC_MAX
    rng = np.random.default_rng(3)
    for sel in (None, rng.permutation(nele)[:nele // 2] + 1, np.array([nele, 1, 1, 7])):
        n = nele if sel is None else sel.size
        ref = _call(out, "mh_beam_gauge_derivs", it, sel, n)
        got = _call(out, "mh_beam_gauge_derivs_noalloc", it, sel, n)
        assert np.isfinite(ref).all()
        assert np.array_equal(got, ref)


def test_noalloc_two_calls_second_smaller():
    out, model, it = _beam_context(5, 4, 3, seed=12)
    nele = model.eletypes[it].nele
    big = np.arange(1, nele + 1)
    small = np.arange(nele, nele - 9, -1)
    first = _call(out, "mh_beam_gauge_derivs_noalloc", it, big, big.size)
    second = _call(out, "mh_beam_gauge_derivs_noalloc", it, small, small.size)
    assert np.array_equal(first, _call(out, "mh_beam_gauge_derivs", it, big, big.size))
    assert np.array_equal(second, _call(out, "mh_beam_gauge_derivs", it, small, small.size))
    assert np.array_equal(second, first[small - 1])


def test_noalloc_refuses_beyond_scratch():
    from muscade_amd import _lib
    out, model, it = _beam_context(3, 3, 2, seed=13)
    sel = np.ones(NOALLOC_MAX + 1, np.int64)
    res = np.zeros((sel.size, GAUGE_OUT))
    r = _lib.lib.mh_beam_gauge_derivs_noalloc(out.ctx.h, out.groups[it], sel.size, _lib.i64ptr(sel), _lib.dptr(res))
    assert r == _lib.MH_ERR_SIZE
    assert "at most 4096" in _lib.lib.mh_last_error(out.ctx.h).decode()
    bad = np.array([0], np.int64)
    r = _lib.lib.mh_beam_gauge_derivs_noalloc(out.ctx.h, out.groups[it], 1, _lib.i64ptr(bad), _lib.dptr(res))
    assert r == _lib.MH_ERR_ARG
