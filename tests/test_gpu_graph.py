"""Whole solver steps captured in a HIP graph (nonlinear_solver.capture /
replay): replays equal eager steps bit for bit, with eager work and
synchronisations between the replays, and the cases whose host-side work
would not replay are refused."""
import pytest
import torch

from tests.cases import case_2d2v, case_2d3v, hip_solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,stepper", [("fp64", "SSP-RK3"), ("fp32-bulk", "SSP-RK3"),
                                               ("fp64", "RK2")])
def test_graph_replays_equal_eager_steps(gpu, precision, stepper):
    c = case_2d2v(seed=3, N_q=64, N_p=12, precision=precision, stepper=stepper)
    eager, graph = hip_solver(c), hip_solver(c)
    graph.capture(2, c.dt)
    assert torch.equal(eager._f, graph._f)  # the capture itself runs nothing
    for _ in range(3):
        eager.strang_timestep(c.dt)
        eager.strang_timestep(c.dt)
        graph.replay()
        torch.ones(1024, device=gpu).cumsum(0)  # eager work between the replays
        torch.cuda.synchronize()
        assert torch.equal(eager._f, graph._f) and torch.equal(eager._mom, graph._mom)
        assert eager.time_elapsed == graph.time_elapsed
    graph.check_divergence()


def test_capture_refuses_fields(gpu):
    c = case_2d3v(seed=0, N_q=8, N_p=6)
    s = hip_solver(c)
    with pytest.raises(NotImplementedError):
        s.capture(1, c.dt)
