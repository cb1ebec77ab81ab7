"""Engine configuration checks that run without a GPU: a learned-theta
preset must name a trained network and use the AABB test; it never runs
fixed theta silently."""
import pytest


def test_learned_theta_needs_a_network():
    from src.core.engine import SimulationEngine
    from src.utils import config as cfg
    with pytest.raises(ValueError, match="no trained network"):
        SimulationEngine(cfg.preset("large_256", THETA_NETWORK=""), state=None)
    with pytest.raises(ValueError, match="MAC='aabb'"):
        SimulationEngine(cfg.preset("large_256", MAC="ref"), state=None)
    with pytest.raises(FileNotFoundError):
        SimulationEngine(cfg.preset("large_256", THETA_NETWORK="no_such_net.npz"), state=None)


def test_learned_presets_name_a_network():
    from src.utils import config as cfg
    for name in ("learned_128", "large_256", "halos_256"):
        c = cfg.preset(name)
        assert c.LEARNED_THETA and c.MAC == "aabb" and c.THETA_NETWORK
