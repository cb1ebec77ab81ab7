"""Pins the CPU oracle with checks that do not restate its formulas:
closed forms, bit de-interleaving, library routines, brute-force top-down
construction and physical invariants."""
import numpy as np
import pytest

from conftest import rel_rms

# ------------------------------------------------------------------ Morton


def _deinterleave(codes, shift):
    """Collect bits shift, shift+3, ... of each code (21 bits)."""
    out = np.zeros_like(codes)
    for b in range(21):
        out |= ((codes >> np.uint64(3 * b + shift)) & np.uint64(1)) << np.uint64(b)
    return out


def test_morton_full21_round_trip(orc):
    rng = np.random.default_rng(1)
    ijk = rng.integers(0, 2 ** 21, size=(3, 5000), dtype=np.int64)
    pos = [(v.astype(np.float64) / 2 ** 21).astype(np.float32) for v in ijk]
    codes, idx = orc.compute_codes(*pos, mask_set=orc.MASK_FULL21)
    for axis in range(3):
        assert np.array_equal(_deinterleave(codes, axis), ijk[axis].astype(np.uint64))
    assert np.array_equal(idx, np.arange(5000, dtype=np.uint32))
    assert int(codes.max()) < 2 ** 63


def test_morton_known_answers(orc):
    z = np.zeros(4, np.float32)
    x = np.array([0.0, 2.0 ** -21, 0.5, 1.0], np.float32)
    c, _ = orc.compute_codes(x, z, z, mask_set=orc.MASK_FULL21)
    # 2^-21 -> ix=1 -> bit 0; 0.5 -> ix=2^20 -> bit 60; 1.0 -> 2^21 wraps to 0
    assert [int(v) for v in c] == [0, 1, 1 << 60, 0]
    c, _ = orc.compute_codes(z, x, z, mask_set=orc.MASK_FULL21)
    assert [int(v) for v in c] == [0, 2, 1 << 61, 0]
    c, _ = orc.compute_codes(z, z, x, mask_set=orc.MASK_FULL21)
    assert [int(v) for v in c] == [0, 4, 1 << 62, 0]
    # negative and NaN saturate to 0
    bad = np.array([-0.25, np.nan, 0.0, 0.0], np.float32)
    c, _ = orc.compute_codes(bad, z, z, mask_set=orc.MASK_FULL21)
    assert [int(v) for v in c] == [0, 0, 0, 0]


def test_morton_ref10_keeps_low_ten_bits(orc):
    """morton.py:33-36 as written == the 21-bit spread of (coordinate & 1023)."""
    rng = np.random.default_rng(2)
    ijk = rng.integers(0, 2 ** 21, size=(3, 5000), dtype=np.int64)
    pos = [(v.astype(np.float64) / 2 ** 21).astype(np.float32) for v in ijk]
    low = [((v & 1023).astype(np.float64) / 2 ** 21).astype(np.float32) for v in ijk]
    c_ref10, _ = orc.compute_codes(*pos, mask_set=orc.MASK_REF10)
    c_low, _ = orc.compute_codes(*low, mask_set=orc.MASK_FULL21)
    assert np.array_equal(c_ref10, c_low)
    assert int(c_ref10.max()) < 2 ** 30


# -------------------------------------------------------------------- sort


def test_argsort_matches_numpy_stable(orc):
    rng = np.random.default_rng(3)
    for n in (1, 7, 1000, 50000):
        codes = rng.integers(0, 50, size=n).astype(np.uint64)  # many ties
        assert np.array_equal(orc.argsort_stable(codes),
                              np.argsort(codes, kind="stable").astype(np.uint32))
        codes = rng.integers(0, 2 ** 63, size=n, dtype=np.int64).astype(np.uint64)
        assert np.array_equal(orc.argsort_stable(codes),
                              np.argsort(codes, kind="stable").astype(np.uint32))


# -------------------------------------------------------------------- tree


def _clz(v):
    return 64 - int(v).bit_length()


def brute_force_children(codes):
    """Top-down radix split of sorted unique codes (Karras 2012 definition):
    a range [lo, hi] splits after the last index whose code shares more than
    clz(c[lo]^c[hi]) leading bits with c[lo]; child ranges of length one are
    leaves, others are internal node gamma (left) / gamma+1 (right)."""
    n = len(codes)
    children = np.full((2 * n - 1, 2), -1, np.int64)
    stack = [(0, n - 1, n)]
    while stack:
        lo, hi, node = stack.pop()
        common = _clz(int(codes[lo]) ^ int(codes[hi]))
        g = lo
        while g + 1 <= hi and _clz(int(codes[lo]) ^ int(codes[g + 1])) > common:
            g += 1
        left = g if g == lo else g + n
        right = g + 1 if g + 1 == hi else g + 1 + n
        children[node] = (left, right)
        if left >= n:
            stack.append((lo, g, left))
        if right >= n:
            stack.append((g + 1, hi, right))
    return children


def test_build_tree_matches_brute_force(orc):
    rng = np.random.default_rng(4)
    for n in (2, 3, 5, 17, 256, 1000):
        codes = np.unique(rng.integers(0, 2 ** 63, size=n, dtype=np.int64).astype(np.uint64))
        while len(codes) < n:
            codes = np.unique(np.concatenate([codes, rng.integers(0, 2 ** 63, size=n, dtype=np.int64).astype(np.uint64)]))[:n]
        children, parents = orc.build_tree(codes)
        assert np.array_equal(children.astype(np.int64), brute_force_children(codes))
        assert orc.validate_tree(children, parents, n) == 0
        assert orc.find_root(parents) == n


def test_build_tree_clustered_codes(orc, ics16):
    codes, _ = orc.compute_codes(*ics16, mask_set=orc.MASK_FULL21)
    codes = np.sort(codes)
    assert len(np.unique(codes)) == len(codes)
    children, parents = orc.build_tree(codes)
    assert np.array_equal(children.astype(np.int64), brute_force_children(codes))
    assert orc.validate_tree(children, parents, len(codes)) == 0


def test_equal_code_runs(orc):
    """With the reference's tie-break (tree_kernels.py:55-56) pairs of equal
    codes still give well-formed trees, while a run of 4 equal codes never
    does; the validator flags the latter."""
    base = np.arange(0, 256, 4, dtype=np.uint64) << np.uint64(10)
    for pos in range(60):
        codes = np.sort(np.concatenate([base, base[pos:pos + 1]]))
        ch, par = orc.build_tree(codes)
        assert orc.validate_tree(ch, par, len(codes)) == 0
        codes = np.sort(np.concatenate([base, np.full(3, base[pos], np.uint64)]))
        ch, par = orc.build_tree(codes)
        assert orc.validate_tree(ch, par, len(codes)) != 0
    codes = np.full(4, 7, np.uint64)
    ch, par = orc.build_tree(codes)
    assert orc.validate_tree(ch, par, 4) != 0


# ---------------------------------------------------------------- upward


def _recursive_multipoles(node, children, pos, mass, out):
    n = len(mass)
    if node < n:
        m = float(mass[node])
        out[node] = (m, [float(pos[k][node]) * m for k in range(3)],
                     [pos[k][node] for k in range(3)], [pos[k][node] for k in range(3)])
        return out[node]
    l, r = children[node]
    L = _recursive_multipoles(l, children, pos, mass, out)
    R = _recursive_multipoles(r, children, pos, mass, out)
    res = (L[0] + R[0], [L[1][k] + R[1][k] for k in range(3)],
           [min(L[2][k], R[2][k]) for k in range(3)], [max(L[3][k], R[3][k]) for k in range(3)])
    out[node] = res
    return res


def test_multipoles_recursive_and_invariants(orc, ics16):
    import sys
    sys.setrecursionlimit(10000)
    codes, _ = orc.compute_codes(*ics16)
    perm = orc.argsort_stable(codes)
    pos = [p[perm] for p in ics16]
    n = len(pos[0])
    mass = np.full(n, 1.0 / n, np.float32)
    children, parents = orc.build_tree(codes[perm])
    nm, nc, nmin, nmax = orc.compute_multipoles(*pos, mass, children, parents)
    out = {}
    _recursive_multipoles(n, children, pos, mass, out)
    for node in range(2 * n - 1):
        m, c, lo, hi = out[node]
        assert nm[node] == m and list(nc[node]) == c
        assert list(nmin[node]) == lo and list(nmax[node]) == hi
    assert abs(nm[n] - float(np.sum(mass.astype(np.float64)))) < 1e-15
    com = orc.normalize_com(nm, nc)
    mean = np.array([np.mean(p.astype(np.float64)) for p in pos])
    assert np.allclose(com[n], mean, rtol=0, atol=1e-12)
    assert np.array_equal(nmin[n], [p.min() for p in pos])
    assert np.array_equal(nmax[n], [p.max() for p in pos])
    # leaves: CoM = position exactly after normalisation
    assert np.array_equal(com[:n].astype(np.float32), np.stack(pos, 1))


def test_top_tree_cache_bfs(orc, ics16):
    codes, _ = orc.compute_codes(*ics16)
    children, parents = orc.build_tree(np.sort(codes))
    n = len(codes)
    top, n2c = orc.build_top_tree_cache(children, n, 256)
    q, order = [n], []
    while q and len(order) < 256:
        v = q.pop(0)
        order.append(v)
        for c in children[v]:
            if c != -1:
                q.append(int(c))
    assert list(top[:len(order)]) == order[:256]
    for slot, v in enumerate(top):
        if v != -1:
            assert n2c[v] == slot
    assert int((n2c != -1).sum()) == int((top != -1).sum())


# ---------------------------------------------------------------- forces


def _tree(orc, pos, mass=None):
    codes, _ = orc.compute_codes(*pos)
    perm = orc.argsort_stable(codes)
    pos = [p[perm] for p in pos]
    n = len(pos[0])
    mass = np.full(n, 1.0 / n, np.float32) if mass is None else mass[perm]
    children, parents = orc.build_tree(codes[perm])
    nm, nc, nmin, nmax = orc.compute_multipoles(*pos, mass, children, parents)
    com = orc.normalize_com(nm, nc)
    return pos, mass, children, nm, com, nmin, nmax


def test_two_body_closed_form(orc):
    pos = [np.array([0.30, 0.40], np.float32), np.array([0.5, 0.5], np.float32),
           np.array([0.5, 0.5], np.float32)]
    mass = np.array([0.25, 0.75], np.float32)
    soft = 0.01
    p, m, ch, nm, com, nmin, nmax = _tree(orc, pos, mass)
    for prec in (64, 32):
        f = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 0.5, 1.0, soft, 1.0, 2,
                               precision=prec)
        d = float(p[0][1]) - float(p[0][0])
        for i, j in ((0, 1), (1, 0)):
            sgn = 1 if i == 0 else -1
            expect = float(m[j]) * sgn * d / (d * d + soft * soft) ** 1.5
            assert abs(f[i, 0] - expect) <= 1e-6 * abs(expect)
            assert f[i, 1] == 0 and f[i, 2] == 0


def test_theta_zero_is_direct_sum(orc, ics16):
    p, m, ch, nm, com, nmin, nmax = _tree(orc, ics16)
    n = len(m)
    f, ni, nv, f64 = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 0.0, 1.0, 0.02 / 16, 1.0,
                                        n, precision=64, counts=True, want_f64=True)
    direct = orc.direct_sum(*p, m, np.arange(n), 1.0, 0.02 / 16)
    assert rel_rms(f64, direct) < 1e-12
    assert np.all(ni == n - 1)        # every other particle, self cut by d2 <= 1e-7
    assert np.all(nv == 2 * n - 1)    # every node visited once


def test_bh_error_versus_theta(orc, ics16):
    """Reference test: error falls and work grows as theta shrinks, but an
    error floor remains (x-extent-only test, minimum image of the CoM).  The
    corrected AABB test converges to the direct sum."""
    p, m, ch, nm, com, nmin, nmax = _tree(orc, ics16)
    n = len(m)
    direct = orc.direct_sum(*p, m, np.arange(n), 1.0, 0.02 / 16)
    for mac, floor in ((orc.MAC_REF, None), (orc.MAC_AABB, 1e-2)):
        errs, inter = [], []
        for th in (0.9, 0.7, 0.5, 0.3, 0.2):
            f, ni, nv = orc.compute_forces(*p, ch, nm, com, nmin, nmax, th, 1.0, 0.02 / 16, 1.0,
                                           n, counts=True, mac=mac)
            errs.append(rel_rms(f, direct))
            inter.append(ni.sum())
        assert all(a > b for a, b in zip(errs, errs[1:]))
        assert all(a < b for a, b in zip(inter, inter[1:]))
        if floor is not None:
            assert errs[-1] < floor
        f = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 0.0, 1.0, 0.02 / 16, 1.0, n, mac=mac)
        assert rel_rms(f, direct) < 1e-6


def test_bucket_rule(orc, ics16):
    """Leaf buckets (AABB test): when every node is rejected the walk reaches
    every leaf in leaf order, so buckets change no bit of the sums, in either
    precision, and no interaction count (only the visits of the internal
    nodes they skip); a whole-tree bucket is the direct sum; and at a real
    theta the buckets can only add exact leaf terms, so the error against
    the direct sum does not grow."""
    p, m, ch, nm, com, nmin, nmax = _tree(orc, ics16)
    n = len(m)
    soft = 0.02 / 16
    for prec in (64, 32):
        base = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 1e-6, 1.0, soft, 1.0, n,
                                  precision=prec, mac=orc.MAC_AABB, counts=True)
        for b in (2, 16, 64):
            got = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 1e-6, 1.0, soft, 1.0, n,
                                     precision=prec, mac=orc.MAC_AABB, counts=True, bucket=b)
            assert np.array_equal(got[0], base[0])
            assert np.array_equal(got[1], base[1])
            # a bucket of two leaves holds no internal node to skip
            assert np.all(got[2] <= base[2]) and (b == 2 or np.all(got[2] < base[2]))
    # the reference test ignores buckets
    ref = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 0.5, 1.0, soft, 1.0, n, precision=32)
    assert np.array_equal(ref, orc.compute_forces(*p, ch, nm, com, nmin, nmax, 0.5, 1.0, soft,
                                                  1.0, n, precision=32, bucket=16))
    direct = orc.direct_sum(*p, m, np.arange(n), 1.0, soft)
    f0 = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 0.5, 1.0, soft, 1.0, n,
                            mac=orc.MAC_AABB)
    f16 = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 0.5, 1.0, soft, 1.0, n,
                             mac=orc.MAC_AABB, bucket=16)
    assert rel_rms(f16, direct) <= rel_rms(f0, direct)
    # 50 particles, one bucket holding the whole tree: the root straddles the
    # minimum-image boundary of every particle, so it is rejected and summed
    rng = np.random.default_rng(8)
    pos = [rng.uniform(0, 1, 50).astype(np.float32) for _ in range(3)]
    p, m, ch, nm, com, nmin, nmax = _tree(orc, pos)
    f, ni, nv, f64 = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 0.5, 1.0, 1e-3, 1.0, 50,
                                        precision=64, mac=orc.MAC_AABB, counts=True,
                                        want_f64=True, bucket=64)
    assert rel_rms(f64, orc.direct_sum(*p, m, np.arange(50), 1.0, 1e-3)) < 1e-12
    assert np.all(ni == 49) and np.all(nv == 51)


def test_straddle_tolerance(orc, ics16):
    """The straddle limit of the AABB test: 0 is the exact guard; a limit
    below the lightest internal node changes nothing; raising it lets heavier
    nodes across the minimum-image boundary through (fewer interactions,
    larger error against the direct sum)."""
    p, m, ch, nm, com, nmin, nmax = _tree(orc, ics16)
    n = len(m)
    soft = 0.02 / 16
    direct = orc.direct_sum(*p, m, np.arange(n), 1.0, soft)
    base = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 0.7, 1.0, soft, 1.0, n, mac=1,
                              counts=True, precision=32)
    tiny = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 0.7, 1.0, soft, 1.0, n, mac=1,
                              counts=True, precision=32, straddle_mass=1.9 / n)
    assert np.array_equal(base[0], tiny[0]) and np.array_equal(base[1], tiny[1])
    errs, inter = [], []
    for mu in (0.0, 1e-3, 1e-2, 1e-1):
        f, ni, nv = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 0.7, 1.0, soft, 1.0, n,
                                       mac=1, counts=True, precision=32, straddle_mass=mu)
        errs.append(rel_rms(f, direct))
        inter.append(int(ni.sum()))
    assert all(a <= b for a, b in zip(errs, errs[1:])) and errs[-1] > 2 * errs[0]
    assert all(a >= b for a, b in zip(inter, inter[1:])) and inter[-1] < inter[0]


@pytest.mark.parametrize("mac", [0, 1])
def test_fp32_mode_tracks_fp64(orc, ics32, mac):
    p, m, ch, nm, com, nmin, nmax = _tree(orc, ics32)
    n = len(m)
    f64, n64, _ = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 0.5, 1.0, 0.02 / 32, 1.0, n,
                                     precision=64, counts=True, mac=mac)
    f32, n32, _ = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 0.5, 1.0, 0.02 / 32, 1.0, n,
                                     precision=32, counts=True, mac=mac)
    # reference test: FP32 tracks FP64 to 1e-5; the AABB test has more
    # borderline acceptance flips between the two precisions
    assert rel_rms(f32, f64) < (1e-5 if mac == 0 else 1e-3)
    assert np.mean(n32 == n64) > (0.999 if mac == 0 else 0.995)


def _code_chain(m):
    """Positions whose FULL21 Morton codes are 2^k for k < m, plus two with
    code 0 (at the origin and 1e-9 from it)."""
    pts = np.zeros((3, m + 2), np.float32)
    for k in range(m):
        pts[k % 3, k + 1] = np.float32(2.0 ** (k // 3 - 21))
    pts[0, m + 1] = np.float32(1e-9)
    return list(pts)


def test_stack_overflow_drops_children(orc):
    """Codes 2^k (k < m) and a pair of zeros give a valid tree (a pair of
    equal codes is split by the tie-break) that is one chain of left (c1)
    edges ending in the pair.  At theta = 0 every internal node is opened.
    With m = 63 the pair's parent sits at stack slot 63, where the reference
    drops children (physics_kernels.py:143): the particles that reach it
    lose exactly the two interactions with the pair.  With m = 40 the stack
    stays below 64 and every particle interacts with every other one beyond
    the 1e-7 cut, as in the direct sum.  The oracle's drop counter reads
    exactly one dropped node per particle (the pair's parent) at m = 63 and
    none at m = 40, for both acceptance tests."""
    for m, drops in ((63, True), (40, False)):
        pos = _code_chain(m)
        codes, _ = orc.compute_codes(*pos)
        perm = orc.argsort_stable(codes)
        assert orc.validate_tree(*orc.build_tree(codes[perm]), m + 2) == 0
        p, ms, ch, nm, com, nmin, nmax = _tree(orc, pos)
        f, ni, nv = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 0.0, 1.0, 1e-3, 1.0, m + 2,
                                       precision=64, counts=True)
        P = np.stack(p, 1).astype(np.float64)
        d = P[None, :, :] - P[:, None, :]
        d -= np.rint(d)
        beyond = ((d ** 2).sum(2) > 1e-7).sum(1)
        if drops:
            lost = beyond - ni
            assert set(np.unique(lost)) == {0, 2} and np.count_nonzero(lost) > 10
        else:
            assert np.array_equal(ni, beyond)
        for mac in (0, 1):
            _, ni2, _, nd = orc.compute_forces(*p, ch, nm, com, nmin, nmax, 0.0, 1.0, 1e-3, 1.0,
                                               m + 2, precision=64, mac=mac, counts=True,
                                               drops=True)
            assert np.array_equal(nd, np.full(m + 2, 1 if drops else 0))
            if mac == 0:
                assert np.array_equal(ni2, ni)


def test_direct_sum_three_body(orc):
    pos = [np.array([0.1, 0.2, 0.9], np.float32), np.array([0.5, 0.6, 0.5], np.float32),
           np.array([0.5, 0.5, 0.5], np.float32)]
    mass = np.array([1.0, 2.0, 3.0], np.float32)
    soft = 0.05
    acc = orc.direct_sum(*pos, mass, np.arange(3), 1.0, soft)
    P = np.stack([a.astype(np.float64) for a in pos], 1)
    for i in range(3):
        e = np.zeros(3)
        for j in range(3):
            if i == j:
                continue
            d = P[j] - P[i]
            d -= np.round(d)  # minimum image: particle 2 sits 0.2 left of 0 via wrap
            e += mass[j] * d / (d @ d + soft * soft) ** 1.5
        assert np.allclose(acc[i], e, rtol=1e-13, atol=0)


# ---------------------------------------------------------------- update


def test_kick_drift_closed_form(orc):
    rng = np.random.default_rng(5)
    n = 4096
    pos = [rng.random(n, dtype=np.float32) for _ in range(3)]
    vel = [rng.standard_normal(n).astype(np.float32) for _ in range(3)]
    force = rng.standard_normal((n, 3)).astype(np.float32)
    a, H, dt = 0.02, 23970.0, 6.5e-7
    out = orc.integrate(*pos, *vel, force, a, H, dt)
    cd = np.float32(1.0 - H * dt)
    ck = np.float32(dt / a)
    for k in range(3):
        v = vel[k] * cd + force[:, k] * ck
        x = pos[k] + v * ck
        x = x - np.floor(x)
        assert np.array_equal(out[3 + k], v)
        assert np.array_equal(out[k], x)
        assert np.all((out[k] >= 0) & (out[k] <= 1))


# ---------------------------------------------------------------- render


def test_render_centre_pixel_and_totals(orc):
    from src.utils import config as cfg
    from src.utils.visualizer import get_mvp_matrix
    c = cfg.preset("repo_128")
    mvp = get_mvp_matrix(0, (0.5, 0.5, 0.5), 1.0, c)
    res = 256
    one = np.array([0.5], np.float32)
    g = orc.render_density(one, one, one, mvp, res)
    ys, xs = np.nonzero(g)
    assert len(ys) == 1 and abs(int(ys[0]) - res // 2) <= 1 and abs(int(xs[0]) - res // 2) <= 1
    # weight 2.5 / w with w = eye-to-centre distance sqrt(1.8^2 + 0.2^2)
    w = float(mvp[3] @ np.array([0.5, 0.5, 0.5, 1.0], np.float32))
    assert abs(w - np.hypot(1.8, 0.2)) < 1e-5
    assert int(g.sum()) == int(np.rint(np.float32(2.5) / np.float32(w) * np.float32(2 ** 20)))
    # total over many particles = sum of individual quantised weights
    rng = np.random.default_rng(6)
    p = [rng.random(20000, dtype=np.float32) for _ in range(3)]
    g = orc.render_density(*p, mvp, res)
    P = np.stack(p + [np.ones(20000, np.float32)], 1)
    clip = (P @ mvp.T.astype(np.float64))
    inside = (clip[:, 3] > 0.01)
    sx = (clip[:, 0] / clip[:, 3] + 1) * 0.5 * res
    sy = (1 - clip[:, 1] / clip[:, 3]) * 0.5 * res
    inside &= (sx >= 0) & (sx < res) & (sy >= 0) & (sy < res)
    expect = np.sum(np.rint(2.5 / clip[inside, 3] * 2 ** 20))
    assert abs(float(g.sum()) - expect) / expect < 1e-5


def test_colorize_palette_known_answers(orc):
    """visualizer.py:115-171 on chosen pixels: the brightest pixel maps to
    val = 1 (red 1.3 -> 1, green 0 after the last band), zero stays black,
    and a value at band edge T1 gives pure (clipped) blue."""
    grid = np.zeros((4, 4), np.float32)
    grid[0, 0] = 10.0
    top = np.log1p(np.float32(1200.0))
    # a pixel with val = 0.6 exactly at the first band edge: img = 0.6^(1/2.2)
    img = np.float32(0.6) ** np.float32(1 / 2.2)
    grid[1, 1] = np.expm1(img * top) / 120.0
    frame, rgb = orc.colorize_frame(grid)
    assert np.allclose(frame[0, 0], [1.0, 0.0, 0.0], atol=1e-6)
    assert np.all(frame[3, 3] == 0)
    assert abs(frame[1, 1, 2] - min(1.0, 0.8 * 1.3)) < 1e-3 and frame[1, 1, 0] == 0
    assert rgb.dtype == np.uint8 and rgb[0, 0, 0] == 255



# ------------------------------------------- hand-derived known answers

@pytest.mark.parametrize("name", ["x_extent", "clamp", "no_clamp"])
def test_reference_mac_known_answers(orc, name):
    """The reference test on hand-built three-particle trees
    (physics_kernels.py:128-136, derivations in kat_cases.py): the x-extent
    alone decides (a node thin in x but tall in y is accepted where a
    largest-extent test would open it), and an x-extent above box/2 is
    replaced by the box, so such a node is always opened for theta <= 1,
    while an x-extent just below box/2 is kept.  Visits, interactions and
    T's acceleration are compared with the hand-derived values."""
    import kat_cases as K
    case = K.FORCE_CASES[name]
    pos, mass, ch, par = K.tree_arrays(case["points"], case["masses"])
    nm, nc, lo, hi = orc.compute_multipoles(*pos, mass, ch, par)
    com = orc.normalize_com(nm, nc)
    expect = K.expected_force(case)
    other = K.opened_force(case) if case["interactions"] == 1 else None
    for prec, tol in ((64, 1e-12), (32, 2e-6)):
        f, ni, nv, f64 = orc.compute_forces(*pos, ch, nm, com, lo, hi, case["theta"], K.G,
                                            K.SOFT, 1.0, 3, off=0, count=1, precision=prec,
                                            counts=True, want_f64=True)
        assert int(ni[0]) == case["interactions"]
        assert int(nv[0]) == case["visits"]
        got = f64[0] if prec == 64 else f[0].astype(np.float64)
        assert np.linalg.norm(got - expect) <= tol * np.linalg.norm(expect)
        if other is not None:  # an opened node would be far from the answer
            assert np.linalg.norm(other - expect) > 1e-3 * np.linalg.norm(expect)


def test_render_orientation_known_answers(orc):
    """Grid rows and columns against the pinhole projection of the
    reference's camera derived in FP64 (kat_cases.pinhole_pixel): a point
    left of the centre on screen lands at a column < res/2 in the centre's
    row, a point above it at a row < res/2 in the centre's column; a
    transposed or flipped grid fails."""
    import kat_cases as K
    from src.utils import config as cfg
    from src.utils.visualizer import get_mvp_matrix
    c = cfg.preset("repo_128")
    for res in (256, 1024):
        mvp = get_mvp_matrix(0, (0.5, 0.5, 0.5), 1.0, c)
        centre = orc.render_density(*[np.array([0.5], np.float32)] * 3, mvp, res)
        (r0,), (c0,) = np.nonzero(centre)
        for name, p in K.SPLAT_POINTS.items():
            row, col, fr, fc = K.pinhole_pixel(p, res)
            g = orc.render_density(*[np.array([v], np.float32) for v in p], mvp, res)
            ys, xs = np.nonzero(g)
            assert len(ys) == 1
            # away from pixel borders the FP32 pipeline must give the FP64 pixel
            if 1e-3 < fr < 1 - 1e-3:
                assert int(ys[0]) == row, (name, res)
            if 1e-3 < fc < 1 - 1e-3:
                assert int(xs[0]) == col, (name, res)
            if name == "left":
                assert xs[0] < res // 2 - 10 and ys[0] == r0
            if name == "up":
                assert ys[0] < res // 2 - 10 and xs[0] == c0
            if name == "right_down":
                assert xs[0] > res // 2 + 5 and ys[0] > res // 2 + 5
