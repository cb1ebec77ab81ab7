"""The whole step captured in HIP graphs: every replay must equal the eager
step bit for bit (positions, velocities, forces, density grid), with eager
work and synchronisations between replays."""
import numpy as np
import pytest

from conftest import gpu_available, zeldovich

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]


def _engine(n_grid=32, mac="ref", **kw):
    from src.core.engine import SimulationEngine
    from src.core.state import SimulationState
    from src.utils import config as cfg
    c = cfg.preset("small_32", N_GRID=n_grid, SOFTENING=0.02 / n_grid, MAC=mac, RES=256, **kw)
    pos, vel = zeldovich(n_grid)
    st = SimulationState(len(pos[0]), pos, vel)
    return SimulationEngine(c, st)


@pytest.mark.parametrize("mac,longest_first", [("ref", False), ("aabb", False), ("aabb", True)])
def test_graph_replays_match_eager(mac, longest_first):
    """With LONGEST_FIRST the captured step also orders the groups by the
    previous launch's pops and stores the new ones."""
    import torch
    eager = _engine(mac=mac, LONGEST_FIRST=longest_first)
    graph = _engine(mac=mac, LONGEST_FIRST=longest_first)
    assert graph.longest_first == longest_first
    steps = 5
    graph.capture_graphs(steps)
    for k in range(steps):
        eager.step()
        eager.camera.capture(eager.steps_done - 1, *eager.state.pos)
        torch.cuda.synchronize()
        graph.step_graph()
        torch.cuda.synchronize()
        # unrelated eager work between replays
        junk = torch.randn(1000, device="cuda").sum()
        torch.cuda.synchronize()
        del junk
        for a, b in zip(eager.state.pos + eager.state.vel, graph.state.pos + graph.state.vel):
            assert torch.equal(a, b), f"step {k}"
        assert torch.equal(eager.state.force, graph.state.force), f"step {k}"
        assert torch.equal(eager.camera.grid_d, graph.camera.grid_d), f"step {k}"
        assert eager.a == graph.a
    with pytest.raises(RuntimeError):
        graph.step_graph()


@pytest.mark.parametrize("longest_first", [False, True])
def test_graph_with_learned_theta_matches_eager(longest_first):
    """The learned per-group theta (features from the step's tree and scale
    factor, MLP, selection) inside the captured step: replays equal eager
    steps bit for bit, theta_group included.  With LONGEST_FIRST the
    captured step also orders the groups by the previous launch's pops
    scaled by the change of their candidates (the two parities' graphs
    swap the candidate buffers as eager steps do)."""
    import torch

    from src.core.learned_theta import DEFAULT_THETAS, LearnedTheta, ThetaModel
    cost = np.linspace(4000.0, 500.0, len(DEFAULT_THETAS))
    engines = []
    for _ in range(2):
        e = _engine(mac="aabb", LONGEST_FIRST=longest_first)
        e.theta_selector = LearnedTheta(ThetaModel(seed=3), DEFAULT_THETAS, cost, 50.0, 0.5)
        engines.append(e)
    eager, graph = engines
    steps = 4
    graph.capture_graphs(steps)
    seen = []
    for k in range(steps):
        eager.step()
        eager.camera.capture(eager.steps_done - 1, *eager.state.pos)
        graph.step_graph()
        torch.cuda.synchronize()
        assert torch.equal(eager.theta_group, graph.theta_group), f"step {k}"
        for a, b in zip(eager.state.pos + eager.state.vel, graph.state.pos + graph.state.vel):
            assert torch.equal(a, b), f"step {k}"
        assert torch.equal(eager.state.force, graph.state.force), f"step {k}"
        assert torch.equal(eager.camera.grid_d, graph.camera.grid_d), f"step {k}"
        seen.append(eager.theta_group.clone())
    assert all(bool(torch.isfinite(t).all()) for t in seen)
