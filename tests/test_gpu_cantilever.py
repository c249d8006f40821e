"""The static cantilever workload (100 EulerBeam3D, L = 10 m, E = 2.1e11 Pa, A = 1e-2 m², I = 1e-5 m⁴,
clamped by 6 Holds, tip load along z ramped to 5e4 N in 10 steps, Newton tolerance 1e-10) with
the device assembler:

- at every converged step, the device's assembled Lλ and Lλx against the oracle's at the same
  state, within 1e-12 of the largest entry, pattern bit-exact;
- the tip trajectory against the one the same loop computes with the oracle's assembly, within
  1e-9 (the Newton tolerance bounds the difference of the two iterations);
- the tip against the large-deflection elastica at every step (tests/test_known_answers.py);
- the 45° bend of examples/StaticBeamAnalysis.jl against Longva and Crisfield."""
import numpy as np
import pytest

from tests.helpers import OracleAssembly, oracle_assemble_sweepx0, relerr
from tests.test_known_answers import bend_tips, cantilever_steps, check_bend, check_elastica, tip

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


@pytest.fixture(scope="module")
def device_run(mc):
    return cantilever_steps(mc, mc.AssemblySweepX)


def test_cantilever_c0_assembly_every_step(mc, device_run):
    model, states = device_run
    out = mc.AssemblySweepX(0, model, states[0].dis)
    for s in states:
        Lλ, nz = out.assemble(s, "iter")
        Re, _ = out.element_blocks(0)
        ref = oracle_assemble_sweepx0(model, s)
        assert np.array_equal(out.colptr, ref["colptr"]) and np.array_equal(out.rowval, ref["rowval"])
        # at equilibrium Lλ sums element forces to nearly zero: its error is measured against the
        # element forces it sums, plus the rounding of an axial force, EA·ε_mach (the strain is a
        # difference of lengths, each known to ε_mach), about 5e-7 N here
        EA = 2.1e11 * 1e-2
        tol = 1e-12 * np.abs(Re).max() + 4 * np.finfo(float).eps * EA
        assert np.abs(Lλ - ref["rhs"]).max() <= tol
        assert relerr(nz, ref["nzval"]) <= 1e-12


def test_cantilever_c0_tip_matches_oracle_loop(mc, device_run):
    _, states = device_run
    _, ref = cantilever_steps(mc, OracleAssembly)
    for s, r in zip(states, ref):
        assert np.abs(tip(mc, s) - tip(mc, r)).max() <= 1e-9 * np.abs(tip(mc, r)).max()


def test_cantilever_c0_elastica_device(mc, device_run):
    check_elastica(mc, device_run[1])


def test_45_degree_bend_device(mc):
    check_bend(bend_tips(mc, mc.AssemblySweepX))
