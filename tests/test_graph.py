"""The whole step captured in HIP graphs: every replay must equal the eager
step bit for bit (positions, velocities, forces, density grid), with eager
work and synchronisations between replays."""
import numpy as np
import pytest

from conftest import gpu_available, zeldovich

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]


def _engine(n_grid=32, mac="ref"):
    from src.core.engine import SimulationEngine
    from src.core.state import SimulationState
    from src.utils import config as cfg
    c = cfg.preset("small_32", N_GRID=n_grid, SOFTENING=0.02 / n_grid, MAC=mac, RES=256)
    pos, vel = zeldovich(n_grid)
    st = SimulationState(len(pos[0]), pos, vel)
    return SimulationEngine(c, st)


@pytest.mark.parametrize("mac", ["ref", "aabb"])
def test_graph_replays_match_eager(mac):
    import torch
    eager = _engine(mac=mac)
    graph = _engine(mac=mac)
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
