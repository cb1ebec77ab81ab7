"""C1 at full size: the 20 steps of repo_128 (N = 128^3 = 2,097,152
Zel'dovich ICs, theta = 0.5, 1024^2 splat) on the GPU, each stage compared
with the oracle on the same input: codes, permutation, children, parents,
FP64 node moments, kick-drift and the fixed-point density grid bit for bit;
accelerations within 1e-5 relative RMS of the reference method's FP64
arithmetic on 16384 targets drawn uniformly without replacement (PCG64
seed 1234).  This is the size at which the upward pass's device-wide climb
spans the most chunks."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]


def test_c1_steps_bit_exact_128(orc):
    import torch

    from src.core.engine import SimulationEngine
    from src.core.state import SimulationState
    from src.utils import config as cfg
    from src.utils.cosmology import generate_zeldovich_ics
    from step_check import check_steps
    c = cfg.preset("repo_128")
    pos, vel = generate_zeldovich_ics(c)
    n = len(pos[0])
    assert n == 128 ** 3
    torch.cuda.set_device(0)
    eng = SimulationEngine(c, SimulationState(n, pos, vel))
    targets = np.sort(np.random.default_rng(1234).choice(n, 16384, replace=False))
    errs = check_steps(orc, eng, c.RUN_STEPS, targets=targets,
                       log=lambda m: print(m, flush=True))
    assert len(errs) == 20
    # the device-wide climb left 2 on every internal node's arrival counter
    cnt = eng.tree_manager.counters.cpu().numpy()
    assert (cnt[:n] == 0).all() and (cnt[n:] == 2).all()
