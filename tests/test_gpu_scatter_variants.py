"""Every shape of the nz scatter (mh_set_scatter_variant: segments per lane × LDS chunk, plain
or nontemporal loads and stores) sums each nz in the same order, so the assembled Lλx.nzval is
bitwise the same for all 16; and it matches the oracle to 1e-12."""
import ctypes

import numpy as np
import pytest

from tests.helpers import build_lattice, oracle_assemble_sweepx0, relerr

pytestmark = pytest.mark.gpu


def test_all_scatter_shapes_bitwise():
    import muscade_amd as mc
    from muscade_amd import _lib
    model, st = build_lattice(7, 6, 5, seed=12, disp=1e-2, rot=0.1)
    out = mc.AssemblySweepX(0, model, st.dis)
    ref = None
    for v in range(16):
        out.ctx.check(_lib.lib.mh_set_scatter_variant(out.ctx.h, v), f"variant {v}")
        _, nz = out.assemble(st, "iter")
        nz = nz.copy()
        if ref is None:
            ref = nz
            orc = oracle_assemble_sweepx0(model, st)
            assert relerr(nz, orc["nzval"]) <= 1e-12
        assert np.array_equal(nz.view(np.uint64), ref.view(np.uint64)), v
    tv, tms = ctypes.c_int(), (ctypes.c_float * 16)()
    out.ctx.check(_lib.lib.mh_get_tuning(out.ctx.h, ctypes.byref(tv), tms), "tuning")
    assert 0 <= tv.value < 16 and all(float(x) > 0 for x in tms)
