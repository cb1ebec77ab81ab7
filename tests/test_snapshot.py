"""Snapshot round trips (CPU tensors; the format holds no pickled objects)."""
import types

import numpy as np
import pytest


def _engine(n=1000, seed=0, **over):
    import torch
    from src.core.state import SimulationState
    from src.utils import config as cfg
    rng = np.random.default_rng(seed)
    pos = [rng.random(n, dtype=np.float32) for _ in range(3)]
    vel = [rng.standard_normal(n).astype(np.float32) for _ in range(3)]
    st = SimulationState(n, pos, vel, device="cpu")
    st.mass = torch.from_numpy(rng.random(n, dtype=np.float32))
    c = cfg.preset("small_32", **over)
    return types.SimpleNamespace(state=st, config=c, a=0.0345678, steps_done=17)


def test_round_trip_is_bit_exact(tmp_path):
    from src.core import snapshot
    e = _engine()
    p = tmp_path / "s.npz"
    snapshot.save(p, e)
    e2 = _engine(seed=1)
    snapshot.restore(e2, p)
    for a, b in zip(e.state.pos + e.state.vel + (e.state.mass,),
                    e2.state.pos + e2.state.vel + (e2.state.mass,)):
        assert np.array_equal(a.numpy().view(np.uint32), b.numpy().view(np.uint32))
    assert e2.a == e.a and e2.steps_done == 17
    arrays, a, steps, cfgd = snapshot.load(p)
    assert cfgd["NAME"] == "small_32" and cfgd["THETA"] == 0.5


def test_config_mismatch_and_corruption_are_refused(tmp_path):
    from src.core import snapshot
    e = _engine()
    p = tmp_path / "s.npz"
    snapshot.save(p, e)
    other = _engine(THETA=0.7)
    with pytest.raises(ValueError, match="different configuration"):
        snapshot.restore(other, p)
    snapshot.restore(other, p, check_config=False)
    d = dict(np.load(p, allow_pickle=False))
    d["config"] = np.array(str(d["config"]).replace('"THETA": 0.5', '"THETA": 0.6'))
    np.savez(tmp_path / "bad.npz", **d)
    with pytest.raises(ValueError, match="does not match its hash"):
        snapshot.load(tmp_path / "bad.npz")
    small = _engine(n=10)
    with pytest.raises(ValueError, match="particles"):
        snapshot.restore(small, p)
