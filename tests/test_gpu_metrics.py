"""Per-step JSON-lines metrics (src/core/metrics.py) on the GPU: the lines
parse, the stage times add up, the walk's counters agree with the engine's
own health counters, the force error is the FP64 direct-sum check of the
step's forces, and a run with metrics ends in the same state bit for bit as
one without."""
import io
import json

import numpy as np
import pytest

from conftest import gpu_available, zeldovich

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]


def _engine(health=False):
    from src.core.engine import SimulationEngine
    from src.core.state import SimulationState
    from src.utils import config as cfg
    c = cfg.preset("small_32")
    pos, vel = zeldovich(32, c.SEED)
    return SimulationEngine(c, SimulationState(len(pos[0]), pos, vel), health=health), c


def test_metrics_lines_and_counts():
    import torch
    from src.core.metrics import STAGES, StepMetrics
    from src.kernels.physics_kernels import direct_sum
    buf = io.StringIO()
    eng, c = _engine()
    m = StepMetrics(buf, counts=True, error_targets=256, error_every=2)
    eng.run(3, metrics=m)
    lines = [json.loads(x) for x in buf.getvalue().splitlines()]
    assert [r["step"] for r in lines] == [0, 1, 2]
    n = eng.state.n
    for r in lines:
        assert r["n"] == n
        assert set(r["ms"]) == set(STAGES) | {"total"}
        assert all(v > 0 for v in r["ms"].values())
        assert abs(r["ms"]["total"] - sum(r["ms"][s] for s in STAGES)) < 1e-6
        assert r["interactions_per_particle"] > 0
        assert r["visits_per_particle"] >= r["interactions_per_particle"]
        assert r["overflow_drops"] == 0 and r["nonfinite_force_components"] == 0
    assert ["force_error" in r for r in lines] == [True, False, True]
    # the per-step counts add up to the health counters of a separate run of
    # the same steps, and with health on the metrics take their differences
    h, _ = _engine(health=True)
    h.run(3)
    rep = h.health_report()
    assert rep["steps"] == 3
    assert rep["interactions"] == round(sum(r["interactions_per_particle"] for r in lines) * n)
    assert rep["visits"] == round(sum(r["visits_per_particle"] for r in lines) * n)
    buf2 = io.StringIO()
    h2, _ = _engine(health=True)
    h2.run(3, metrics=StepMetrics(buf2, counts=True))
    lines2 = [json.loads(x) for x in buf2.getvalue().splitlines()]
    for r, r2 in zip(lines, lines2):
        assert r2["interactions_per_particle"] == r["interactions_per_particle"]
        assert r2["visits_per_particle"] == r["visits_per_particle"]
    assert h2.health_report()["interactions"] == rep["interactions"]
    # the error of step 0: the FP64 direct-sum check of the first step's forces
    ref_eng, _ = _engine()
    ref_eng.build_tree()
    ref_eng.compute_forces()
    st = ref_eng.state
    tg = torch.from_numpy(np.sort(np.random.default_rng(1234).choice(n, 256, replace=False)))
    tg = tg.to(st.force.device)
    ref = direct_sum(*st.pos, st.mass, tg, c.G, c.SOFTENING)
    got = st.force.index_select(0, tg).double()
    err = float(((got - ref) ** 2).sum().sqrt() / (ref ** 2).sum().sqrt())
    assert abs(lines[0]["force_error"]["rel_rms_vs_fp64_direct_sum"] - err) <= 1e-12 * err
    assert lines[0]["force_error"]["max_rel"] >= err * 0.999


def test_metrics_do_not_change_the_run():
    from src.core.metrics import StepMetrics
    a, _ = _engine()
    b, _ = _engine()
    a.run(3)
    # (stage events and the force check only; counts=True runs the counting
    # walk, another instantiation of the traversal)
    b.run(3, metrics=StepMetrics(io.StringIO(), error_targets=64))
    for x, y in zip(a.state.pos + a.state.vel, b.state.pos + b.state.vel):
        assert np.array_equal(x.cpu().numpy(), y.cpu().numpy())
    assert np.array_equal(a.camera.grid_d.cpu().numpy(), b.camera.grid_d.cpu().numpy())
    c, _ = _engine()
    c.run(1, metrics=StepMetrics(io.StringIO(), counts=True))
    assert c.stats is None  # the counting buffer is handed back after each step
