"""Host-side logic of the learned selection's evaluation (no GPU)."""
import math

import numpy as np


def test_hindsight_choice_meets_the_target_exactly():
    """1000 groups, a cheap candidate (cost 1, squared error g + 1 for group
    g) and an exact one (cost 2, no error).  The target admits the 600
    smallest errors (sum 600 * 601 / 2) and no more, so the cheapest oracle
    choice switches exactly the 400 largest-error groups: mean cost 1.4.  On
    the log-lambda grid alone the switch threshold moves in 1.26x steps and
    the mean cost can land up to a step above that; the bisection finds it."""
    from src.core.learned_theta import hindsight_choice
    G = 1000
    E = np.zeros((G, 2))
    E[:, 0] = np.arange(1, G + 1, dtype=np.float64)
    cost = np.array([1.0, 2.0])
    target = math.sqrt(600 * 601 / 2)
    lam, mc = hindsight_choice(E, cost, 1.0, target)
    assert abs(mc - 1.4) < 1e-12
    k = np.argmin(cost[None, :] + lam * E, axis=1)
    assert int(k.sum()) == 400 and math.sqrt(E[np.arange(G), k].sum()) <= target
    # a target no candidate choice meets, and one the cheapest choice meets
    assert hindsight_choice(E, cost, 1.0, -1.0) is None
    lam, mc = hindsight_choice(E, cost, 1.0, math.sqrt(E[:, 0].sum()))
    assert mc == 1.0
