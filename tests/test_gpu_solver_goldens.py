"""The product's solver loops with the device assembler against the reference's solver goldens:
the reference's test elements run as host residual elements whose R, ∂R/∂δX and ∂R/∂δr the
device scatters (MH_KIND_HOST, mh_set_host_blocks, mh_set_host_dr). The CPU variant
(tests/test_solver_goldens.py) holds the goldens and their sources."""
import pytest

from tests.test_solver_goldens import check_sweepx1, check_sweepx2, sweepx1_decay, sweepx2_sdof

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


@pytest.mark.parametrize("graph", [False, True])
def test_sweepx2_sdof_oscillator_golden_device(mc, graph):
    check_sweepx2(*sweepx2_sdof(mc, mc.AssemblySweepX, graph=graph))


def test_sweepx1_exponential_decay_golden_device(mc):
    check_sweepx1(*sweepx1_decay(mc, mc.AssemblySweepX))


def test_directxua001_solution_golden_device(mc):
    """TestDirectXUA001.jl:33-37 with SweepX{0} on the device (Holds and DofLoads by the device
    kernels, the springs as host residual elements) and DirectXUA's default assembler."""
    from tests.test_solver_goldens import check_directxua001, directxua001
    check_directxua001(directxua001(mc, mc.AssemblySweepX, mc.AssemblyDirect))
