"""GPU parity: every stage of the HIP path against the CPU oracle on the
same seeded inputs (bit-exact for codes, permutation, tree, moments, update
and density grid; forces within the stated tolerances)."""
import numpy as np
import pytest

from conftest import gpu_available, rel_rms, zeldovich

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    return torch


def dev(a, torch):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint64:
        a = a.view(np.int64)
    elif a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).to("cuda")


def host(t, dtype=None):
    a = t.detach().cpu().numpy()
    return a.view(dtype) if dtype is not None else a


def sorted_case(orc, pos, mask_set=0):
    codes, _ = orc.compute_codes(*pos, mask_set=mask_set)
    perm = orc.argsort_stable(codes)
    return [p[perm] for p in pos], codes[perm]


# ----------------------------------------------------------------- Morton

@pytest.mark.parametrize("mask_set", [0, 1])
def test_codes_bit_exact(orc, torch_mod, ics32, mask_set):
    torch = torch_mod
    from src.kernels.morton import compute_codes
    pos = [p.copy() for p in ics32]
    pos[0][:4] = [1.0, -0.0, 0.5, np.nextafter(np.float32(1.0), np.float32(0.0))]
    n = len(pos[0])
    ref, idx = orc.compute_codes(*pos, mask_set=mask_set)
    codes = torch.zeros(n, dtype=torch.int64, device="cuda")
    ind = torch.zeros(n, dtype=torch.int32, device="cuda")
    compute_codes[(n + 255) // 256, 256](*[dev(p, torch) for p in pos], codes, ind, n,
                                         mask_set=mask_set)
    assert np.array_equal(host(codes, np.uint64), ref)
    assert np.array_equal(host(ind, np.uint32), idx)


# ------------------------------------------------------------------- sort

def _gpu_sort(torch, codes_np, bits=63):
    from src.kernels import _hip
    n = len(codes_np)
    ctx = _hip.Context(n)
    c = dev(codes_np, torch)
    out = torch.empty_like(c)
    perm = torch.empty(n, dtype=torch.int32, device="cuda")
    L = _hip.load()
    _hip.check(L.cosmo_sort_by_code(ctx.handle, _hip.ptr(c, _hip.U64, "c"),
                                    _hip.ptr(out, _hip.U64, "o"), _hip.ptr(perm, _hip.U32, "p"),
                                    n, bits, _hip.stream_handle()), "sort")
    torch.cuda.synchronize()
    return host(out, np.uint64), host(perm, np.uint32)


@pytest.mark.parametrize("n", [1, 2, 63, 4095, 4096, 4097, 100003])
def test_sort_matches_stable_argsort(orc, torch_mod, n):
    rng = np.random.default_rng(n)
    for codes in (rng.integers(0, 2 ** 63, size=n, dtype=np.int64).astype(np.uint64),
                  rng.integers(0, 7, size=n).astype(np.uint64) << np.uint64(40)):
        out, perm = _gpu_sort(torch_mod, codes)
        expect = orc.argsort_stable(codes)
        assert np.array_equal(perm, expect)
        assert np.array_equal(perm, np.argsort(codes, kind="stable"))
        assert np.array_equal(out, codes[expect])


def test_sort_ref10_bits_and_c0(orc, torch_mod, ics32):
    for ms, bits in ((0, 63), (1, 30)):
        codes, _ = orc.compute_codes(*ics32, mask_set=ms)
        out, perm = _gpu_sort(torch_mod, codes, bits)
        assert np.array_equal(perm, orc.argsort_stable(codes))


def test_state_sort_buffer_by_morton(orc, torch_mod, ics32):
    torch = torch_mod
    from src.core.state import SimulationState
    from src.utils import config as cfg
    pos, vel = zeldovich(32)
    n = len(pos[0])
    st = SimulationState(n, pos, vel)
    st.mass = dev(np.linspace(0.5, 1.5, n).astype(np.float32), torch)
    m0 = host(st.mass)
    st.sort_buffer_by_morton(cfg.preset("small_32"))
    codes, _ = orc.compute_codes(*pos)
    perm = orc.argsort_stable(codes)
    assert np.array_equal(host(st.sort_idx, np.uint32), perm)
    assert np.array_equal(host(st.codes, np.uint64), codes[perm])
    for a, b in zip(st.pos + st.vel, pos + vel):
        assert np.array_equal(host(a), b[perm])
    assert np.array_equal(host(st.mass), m0[perm])


# ------------------------------------------------------------------- tree

def _gpu_tree(torch, codes_sorted):
    from src.kernels.tree_kernels import build_tree, reset_tree_buffers
    n = len(codes_sorted)
    M = 2 * n - 1
    ch = torch.zeros((M, 2), dtype=torch.int32, device="cuda")
    par = torch.zeros(M, dtype=torch.int32, device="cuda")
    cnt = torch.ones(M, dtype=torch.int32, device="cuda")
    root = torch.ones(1, dtype=torch.int32, device="cuda")
    reset_tree_buffers[(M + 255) // 256, 256](ch, par, cnt, root)
    build_tree[(n + 255) // 256, 256](dev(codes_sorted, torch), ch, par)
    return ch, par, cnt, root


def test_build_tree_bit_exact(orc, torch_mod, ics32):
    torch = torch_mod
    from src.kernels.tree_kernels import find_root, validate_tree
    _, codes = sorted_case(orc, ics32)
    ch, par, cnt, root = _gpu_tree(torch, codes)
    ref_ch, ref_par = orc.build_tree(codes)
    assert np.array_equal(host(ch), ref_ch)
    assert np.array_equal(host(par), ref_par)
    assert int(host(cnt).max()) == 0
    M = 2 * len(codes) - 1
    find_root[(M + 255) // 256, 256](par, root)
    assert int(root.item()) == orc.find_root(ref_par) == len(codes)
    assert validate_tree(ch, par, len(codes)) == 0


def test_build_tree_equal_codes_and_validator(orc, torch_mod, ics16):
    """REF10 codes at 16^3 contain equal-code runs; children follow the
    reference tie-break bit for bit and the validator agrees with the oracle."""
    torch = torch_mod
    from src.kernels.tree_kernels import validate_tree
    _, codes = sorted_case(orc, ics16, mask_set=1)
    ch, par, _, _ = _gpu_tree(torch, codes)
    ref_ch, ref_par = orc.build_tree(codes)
    assert np.array_equal(host(ch), ref_ch)
    ref_flags = orc.validate_tree(ref_ch, ref_par, len(codes))
    gpu_flags = validate_tree(ch, par, len(codes))
    assert (gpu_flags == 0) == (ref_flags == 0)
    dup = np.full(4, 7, np.uint64)
    ch, par, _, _ = _gpu_tree(torch, dup)
    assert validate_tree(ch, par, 4) != 0


def _gpu_moments(torch, pos, mass, ch, par):
    from src.kernels.tree_kernels import compute_multipoles
    n = len(pos[0])
    M = 2 * n - 1
    nm = torch.zeros(M, dtype=torch.float64, device="cuda")
    nc = torch.zeros((M, 3), dtype=torch.float64, device="cuda")
    lo = torch.zeros((M, 3), dtype=torch.float32, device="cuda")
    hi = torch.zeros((M, 3), dtype=torch.float32, device="cuda")
    cnt = torch.zeros(M, dtype=torch.int32, device="cuda")
    compute_multipoles[(n + 255) // 256, 256](*[dev(p, torch) for p in pos], dev(mass, torch),
                                              dev(ch, torch), dev(par, torch), nm, nc, lo, hi,
                                              cnt)
    return nm, nc, lo, hi


def test_multipoles_bit_exact(orc, torch_mod, ics32):
    torch = torch_mod
    from src.kernels.tree_kernels import normalize_com
    pos, codes = sorted_case(orc, ics32)
    n = len(codes)
    mass = np.random.default_rng(7).uniform(0.5, 1.5, n).astype(np.float32) / n
    ch, par = orc.build_tree(codes)
    nm, nc, lo, hi = _gpu_moments(torch, pos, mass, ch, par)
    r_nm, r_nc, r_lo, r_hi = orc.compute_multipoles(*pos, mass, ch, par)
    assert np.array_equal(host(nm), r_nm)
    assert np.array_equal(host(nc), r_nc)
    assert np.array_equal(host(lo), r_lo)
    assert np.array_equal(host(hi), r_hi)
    M = 2 * n - 1
    normalize_com[(M + 255) // 256, 256](nm, nc, M)
    assert np.array_equal(host(nc), orc.normalize_com(r_nm, r_nc))


@pytest.mark.parametrize("n,power", [(1, 1), (2, 1), (1023, 1), (1025, 3), (20001, 3),
                                     (70000, 2)])
def test_multipoles_bit_exact_clustered(orc, torch_mod, n, power):
    """Upward pass on clustered points (u^power per axis) with sizes that do
    not fill the 1024-leaf chunks of the kernel: deep subtrees cross chunk
    boundaries, so both the in-chunk climb and the device-wide climb run.
    Bit-exact moments and counters equal to 2 on every internal node."""
    torch = torch_mod
    rng = np.random.default_rng(n)
    pos, codes = sorted_case(orc, list((rng.random((3, n)) ** power).astype(np.float32)))
    ch, par = orc.build_tree(codes)
    if orc.validate_tree(ch, par, n) != 0:
        pytest.skip("equal Morton codes give a malformed tree for this draw")
    mass = rng.uniform(0.5, 1.5, n).astype(np.float32) / n
    from src.kernels.tree_kernels import compute_multipoles
    M = 2 * n - 1
    nm = torch.zeros(M, dtype=torch.float64, device="cuda")
    nc = torch.zeros((M, 3), dtype=torch.float64, device="cuda")
    lo = torch.zeros((M, 3), dtype=torch.float32, device="cuda")
    hi = torch.zeros((M, 3), dtype=torch.float32, device="cuda")
    cnt = torch.zeros(M, dtype=torch.int32, device="cuda")
    compute_multipoles[(n + 255) // 256, 256](*[dev(p, torch) for p in pos], dev(mass, torch),
                                              dev(ch, torch), dev(par, torch), nm, nc, lo, hi,
                                              cnt)
    r_nm, r_nc, r_lo, r_hi = orc.compute_multipoles(*pos, mass, ch, par)
    assert np.array_equal(host(nm), r_nm)
    assert np.array_equal(host(nc), r_nc)
    assert np.array_equal(host(lo), r_lo)
    assert np.array_equal(host(hi), r_hi)
    c = host(cnt)
    assert (c[:n] == 0).all() and (c[n:] == 2).all()


def test_top_tree_cache(orc, torch_mod, ics16):
    torch = torch_mod
    from src.kernels.tree_kernels import build_top_tree_cache
    _, codes = sorted_case(orc, ics16)
    ch, par = orc.build_tree(codes)
    n = len(codes)
    top = torch.full((256,), -1, dtype=torch.int32, device="cuda")
    n2c = torch.full((2 * n - 1,), -1, dtype=torch.int32, device="cuda")
    root = torch.tensor([n], dtype=torch.int32, device="cuda")
    build_top_tree_cache[1, 1](dev(ch, torch), root, top, n2c, 256)
    r_top, r_n2c = orc.build_top_tree_cache(ch, n, 256)
    assert np.array_equal(host(top), r_top)
    assert np.array_equal(host(n2c), r_n2c)


# ----------------------------------------------------------------- forces

def _case(orc, pos, mass=None):
    pos, codes = sorted_case(orc, pos)
    n = len(codes)
    mass = np.full(n, 1.0 / n, np.float32) if mass is None else mass
    ch, par = orc.build_tree(codes)
    nm, nc, lo, hi = orc.compute_multipoles(*pos, mass, ch, par)
    com = orc.normalize_com(nm, nc)
    return pos, mass, ch, nm, com, lo, hi


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_compute_forces_reference_mac(orc, torch_mod, ics32, theta):
    """compute_forces drop-in (batched launches like engine.py:91-111, which
    take the split walk): <= 1e-5 relative RMS against the reference method's
    FP64 arithmetic.  The split walk sums each particle's terms per top-tree
    subtree, so against the FP32 oracle (one running sum in walk order) the
    two FP32 roundings of the same terms differ by up to about twice the
    oracle's own FP32 error (2e-6 at theta = 0.3): 5e-6.  The visited and
    accepted nodes themselves are compared exactly in
    test_traverse_split_matches_walk."""
    torch = torch_mod
    from src.kernels.physics_kernels import compute_forces
    pos, mass, ch, nm, com, lo, hi = _case(orc, ics32)
    n = len(mass)
    soft = 0.02 / 32
    force = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    root = torch.tensor([n], dtype=torch.int32, device="cuda")
    args = [dev(p, torch) for p in pos] + [dev(ch, torch), dev(nm, torch), dev(com, torch),
                                           dev(lo, torch), dev(hi, torch), force]
    for off in range(0, n, 5000):
        c = min(5000, n - off)
        compute_forces[(c + 255) // 256, 256](*args, theta, 1.0, soft, 1.0, root, None, None, off)
    f = host(force)
    f64 = orc.compute_forces(*pos, ch, nm, com, lo, hi, theta, 1.0, soft, 1.0, n, precision=64)
    f32 = orc.compute_forces(*pos, ch, nm, com, lo, hi, theta, 1.0, soft, 1.0, n, precision=32)
    e64, e32 = rel_rms(f, f64), rel_rms(f, f32)
    print(f"theta={theta} relRMS vs FP64 {e64:.3e} vs FP32 oracle {e32:.3e}")
    assert e64 <= 1e-5
    assert e32 <= 5e-6


def _split_case(orc, torch, pos, theta, soft, slices, mass=None):
    from src.kernels.physics_kernels import traverse, traverse_split
    from src.kernels.tree_kernels import pack_nodes
    pos, mass, ch, nm, com, lo, hi = _case(orc, pos, mass)
    n = len(mass)
    nodes = pack_nodes(dev(ch, torch), dev(nm, torch), dev(com, torch), dev(lo, torch),
                       dev(hi, torch), 1.0, 0)
    root = torch.tensor([n], dtype=torch.int32, device="cuda")
    P = [dev(p, torch) for p in pos]
    f32, ni, nv = orc.compute_forces(*pos, ch, nm, com, lo, hi, theta, 1.0, soft, 1.0, n,
                                     precision=32, mac=0, counts=True)
    for off, count in slices:
        end = min(n, off + count)
        f_one = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
        f_split = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
        st = torch.zeros(2, dtype=torch.int64, device="cuda")
        traverse(*P, nodes, root, f_one, theta, 1.0, soft, 1.0, off=off, count=count)
        traverse_split(*P, nodes, root, f_split, theta, 1.0, soft, 1.0, off=off, count=count,
                       stats=st)
        s = host(st)
        # the same visits and interactions, particle by particle summed
        assert int(s[0]) == int(ni[off:end].sum())
        assert int(s[1]) == int(nv[off:end].sum())
        a, b = host(f_split)[off:end], host(f_one)[off:end]
        # the same terms summed per subtree: FP32 rounding differs from the
        # one running sum by up to about twice that sum's own error
        assert rel_rms(a, f32[off:end]) <= 5e-6
        assert rel_rms(a, b) <= 5e-6
        # nothing outside the slice is written
        assert not host(f_split)[:off].any() and not host(f_split)[end:].any()


@pytest.mark.parametrize("theta", [0.3, 0.7])
def test_traverse_split_matches_walk(orc, torch_mod, ics32, theta):
    """Split walk (one wavefront per top-tree subtree and group): exactly the
    oracle's visits and interactions on slices of every size class, forces
    within 5e-6 of the FP32 oracle and of the one-wavefront walk (the same
    terms, summed per subtree)."""
    _split_case(orc, torch_mod, ics32, theta, 0.02 / 32,
                [(0, 5000), (12345, 777), (0, 32768), (32000, 5000), (100, 1)])


def test_traverse_split_stack_overflow(orc, torch_mod):
    """Codes 2^k (k < 63) and a pair of zeros give a valid tree that is one
    chain of left (c1) edges ending in the pair, whose parent sits at stack
    slot 63: at theta = 0 the walk drops the pair there
    (physics_kernels.py:143).  The split walk, whose top tree covers only
    the first 24 levels, drops the same interactions (the oracle's counts)."""
    from test_oracle_stages import _code_chain
    _split_case(orc, torch_mod, _code_chain(63), 0.0, 1e-3, [(0, 65), (5, 40)])


@pytest.mark.parametrize("mac", [0, 1])
def test_traverse_overflow_drops(orc, torch_mod, mac):
    """The code chain of 63 levels at theta = 0: every particle's walk finds
    the full 64-entry stack at the pair's parent and drops its children
    (physics_kernels.py:143).  The traversal's drop counter equals the
    oracle's per-particle drops summed (one per particle), and its
    interaction and visit counts and forces equal the oracle's."""
    torch = torch_mod
    from test_oracle_stages import _code_chain

    from src.kernels.physics_kernels import traverse
    from src.kernels.tree_kernels import pack_nodes
    pos, mass, ch, nm, com, lo, hi = _case(orc, _code_chain(63))
    n = len(mass)
    nodes = pack_nodes(dev(ch, torch), dev(nm, torch), dev(com, torch), dev(lo, torch),
                       dev(hi, torch), 1.0, mac)
    root = torch.tensor([n], dtype=torch.int32, device="cuda")
    force = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    stats = torch.zeros(4, dtype=torch.int64, device="cuda")
    traverse(*[dev(p, torch) for p in pos], nodes, root, force, 0.0, 1.0, 1e-3, 1.0, mac=mac,
             stats=stats)
    f32, ni, nv, nd = orc.compute_forces(*pos, ch, nm, com, lo, hi, 0.0, 1.0, 1e-3, 1.0, n,
                                         precision=32, mac=mac, counts=True, drops=True)
    s = host(stats)
    assert int(nd.sum()) == n
    assert int(s[3]) == int(nd.sum())
    assert int(s[0]) == int(ni.sum()) and int(s[1]) == int(nv.sum())
    assert rel_rms(host(force), f32) <= 2e-6
    # a stack that never fills drops nothing
    f2 = torch.zeros_like(force)
    st2 = torch.zeros(4, dtype=torch.int64, device="cuda")
    traverse(*[dev(p, torch) for p in pos], nodes, root, f2, 0.5, 1.0, 1e-3, 1.0, mac=mac,
             stats=st2)
    assert int(host(st2)[3]) == int(orc.compute_forces(*pos, ch, nm, com, lo, hi, 0.5, 1.0,
                                                       1e-3, 1.0, n, precision=32, mac=mac,
                                                       drops=True)[1].sum())


@pytest.mark.parametrize("mac", ["ref", "aabb"])
def test_engine_health_counts_nonfinite(torch_mod, mac):
    """Health mode on a clean state, on a state with one NaN mass and on one
    with a mass of 3e38.  A NaN mass makes the CoM of the leaf and of every
    node above it NaN: those nodes fail every comparison, so they are opened
    (internal) or skipped (the leaf), the walk still ends, and no
    acceleration is touched -- the reference method's own outcome, which
    the oracle shows as well.  A mass of 3e38 keeps every CoM finite and
    sends every other particle's acceleration to infinity.  The device
    counter equals the number of non-finite components in the force array
    summed over the two steps, and no stack overflows."""
    torch = torch_mod
    from src.core.engine import SimulationEngine
    from src.core.state import SimulationState
    from src.utils import config as cfg
    from src.utils.cosmology import generate_ics
    c = cfg.preset("small_32", MAC=mac, BUCKET=16 if mac == "aabb" else 0)
    pos, vel = generate_ics(c, 7)
    n = len(pos[0])
    for bad in (None, float("nan"), 3e38):
        st = SimulationState(n, pos, vel)
        if bad is not None:
            st.mass[1234] = bad
        eng = SimulationEngine(c, st, health=True)
        for _ in range(2):
            eng.build_tree()
            eng.compute_forces()
        torch.cuda.synchronize()
        rep = eng.health_report()
        assert rep["steps"] == 2 and rep["overflow_drops"] == 0
        last = int((~torch.isfinite(eng.state.force)).sum().item())
        assert rep["nonfinite_force_components"] == 2 * last, bad
        if bad == 3e38:
            assert last == 3 * (n - 1)
        else:
            assert last == 0
            assert rep["interactions"] > 0 and rep["visits"] >= rep["interactions"]


@pytest.mark.parametrize("mac", [0, 1])
def test_traverse_interaction_counts(orc, torch_mod, ics32, mac):
    torch = torch_mod
    from src.kernels.physics_kernels import traverse
    from src.kernels.tree_kernels import pack_nodes
    pos, mass, ch, nm, com, lo, hi = _case(orc, ics32)
    n = len(mass)
    soft = 0.02 / 32
    nodes = pack_nodes(dev(ch, torch), dev(nm, torch), dev(com, torch), dev(lo, torch),
                       dev(hi, torch), 1.0, mac)
    force = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    stats = torch.zeros(4, dtype=torch.int64, device="cuda")
    root = torch.tensor([n], dtype=torch.int32, device="cuda")
    traverse(*[dev(p, torch) for p in pos], nodes, root, force, 0.5, 1.0, soft, 1.0, mac=mac,
             stats=stats)
    f32, ni, nv = orc.compute_forces(*pos, ch, nm, com, lo, hi, 0.5, 1.0, soft, 1.0, n,
                                     precision=32, mac=mac, counts=True)
    s = host(stats)
    assert int(s[0]) == int(ni.sum())
    assert int(s[1]) == int(nv.sum())
    assert rel_rms(host(force), f32) <= 2e-6
    # per-group counts: each wavefront group's interactions and visits are the
    # oracle's per-particle counts summed over its 64 particles
    W = (n + 63) // 64
    gs = torch.zeros(4 + 3 * W, dtype=torch.int64, device="cuda")
    traverse(*[dev(p, torch) for p in pos], nodes, root, force, 0.5, 1.0, soft, 1.0, mac=mac,
             stats=gs, group_stats=True)
    g = host(gs)
    per = g[4:].reshape(W, 3)
    pad = np.zeros(W * 64 - n, dtype=np.int64)
    assert np.array_equal(per[:, 0], np.concatenate([ni, pad]).reshape(W, 64).sum(1))
    assert np.array_equal(per[:, 1], np.concatenate([nv, pad]).reshape(W, 64).sum(1))
    assert np.array_equal(g[:4], s) and int(per[:, 2].sum()) == int(s[2])


@pytest.mark.parametrize("mac", [0, 1])
def test_traverse_xcd_block_mapping(orc, torch_mod, mac):
    """A launch of 293 blocks of 4 wavefronts: the first 256 blocks go to
    the XCDs in runs of 32 (the remapped part), the last 37 keep their own
    index.  Every group is walked exactly once with the oracle's visits and
    interactions (per group) and forces within FP32 rounding."""
    torch = torch_mod
    from src.kernels.physics_kernels import traverse
    from src.kernels.tree_kernels import pack_nodes
    n = 293 * 256 - 100
    rng = np.random.default_rng(293)
    pos = [rng.uniform(0, 1, n).astype(np.float32) ** 2 for _ in range(3)]
    pos, mass, ch, nm, com, lo, hi = _case(orc, pos)
    soft = 1e-3
    nodes = pack_nodes(dev(ch, torch), dev(nm, torch), dev(com, torch), dev(lo, torch),
                       dev(hi, torch), 1.0, mac)
    W = (n + 63) // 64
    force = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    gs = torch.zeros(4 + 3 * W, dtype=torch.int64, device="cuda")
    root = torch.tensor([n], dtype=torch.int32, device="cuda")
    traverse(*[dev(p, torch) for p in pos], nodes, root, force, 0.5, 1.0, soft, 1.0, mac=mac,
             stats=gs, group_stats=True)
    f32, ni, nv = orc.compute_forces(*pos, ch, nm, com, lo, hi, 0.5, 1.0, soft, 1.0, n,
                                     precision=32, mac=mac, counts=True)
    per = host(gs)[4:].reshape(W, 3)
    pad = np.zeros(W * 64 - n, dtype=np.int64)
    assert np.array_equal(per[:, 0], np.concatenate([ni, pad]).reshape(W, 64).sum(1))
    assert np.array_equal(per[:, 1], np.concatenate([nv, pad]).reshape(W, 64).sum(1))
    assert rel_rms(host(force), f32) <= 2e-6


def _leaf_ranges(ch, n):
    """(first, last) leaf of every internal node from the children array."""
    first = np.arange(2 * n - 1, dtype=np.int64)
    last = first.copy()
    order, stack = [], [n]
    while stack:
        v = stack.pop()
        order.append(v)
        if v >= n:
            stack.extend(ch[v])
    for v in reversed(order):
        if v >= n:
            a, b = ch[v]
            first[v], last[v] = min(first[a], first[b]), max(last[a], last[b])
    return first[n:], last[n:]


@pytest.mark.parametrize("bucket", [2, 16, 64])
def test_traverse_buckets(orc, torch_mod, ics32, bucket):
    """AABB records with leaf buckets: the build's leaf ranges match the
    tree's; the traversal's per-particle interaction and visit counts equal
    the FP32 oracle's bucket rule (totals and per group, fixed and per-group
    theta, plain and compensated sums, exact and tolerant minimum-image
    guard) and the forces agree to FP32 rounding."""
    torch = torch_mod
    from src.kernels.physics_kernels import traverse
    from src.kernels.tree_kernels import build_tree_spans, normalize_pack_buckets
    pos, mass, ch, nm, com, lo, hi = _case(orc, ics32)
    _, codes = sorted_case(orc, ics32)
    n = len(mass)
    soft = 0.02 / 32
    m = 2 * n - 1
    children = torch.full((m, 2), -1, dtype=torch.int32, device="cuda")
    parents = torch.full((m,), -1, dtype=torch.int32, device="cuda")
    far = torch.zeros(n - 1, dtype=torch.int32, device="cuda")
    build_tree_spans(dev(codes, torch), children, parents, far)
    assert np.array_equal(host(children), ch)
    f_ = host(far).astype(np.int64)
    i = np.arange(n - 1)
    first, last = _leaf_ranges(ch, n)
    assert np.array_equal(np.minimum(i, f_), first) and np.array_equal(np.maximum(i, f_), last)
    # the bucket pack normalises the CoM sums itself
    _, csum, _, _ = orc.compute_multipoles(*pos, mass, ch, orc.build_tree(codes)[1])
    nodes = normalize_pack_buckets(children, far, dev(nm, torch), dev(csum, torch),
                                   dev(lo, torch), dev(hi, torch), 1.0, bucket)
    root = torch.tensor([n], dtype=torch.int32, device="cuda")
    P = [dev(p, torch) for p in pos]
    W = (n + 63) // 64
    tg = np.random.default_rng(6).uniform(0.1, 0.9, W).astype(np.float32)
    # (theta, per-group theta, compensated, straddle mass)
    for theta, group_theta, comp, mu in ((0.3, None, False, 0.0), (0.7, None, False, 0.0),
                                         (0.5, tg, False, 0.0), (0.3, None, True, 0.0),
                                         (0.6, None, True, 4e-3), (0.5, tg, False, 1e-3)):
        force = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
        gs = torch.zeros(4 + 3 * W, dtype=torch.int64, device="cuda")
        traverse(*P, nodes, root, force, theta, 1.0, soft, 1.0, mac=1, stats=gs,
                 group_stats=True, compensated=comp, straddle_mass=mu,
                 theta_group=None if group_theta is None else dev(group_theta, torch))
        per_particle = None if group_theta is None else np.repeat(group_theta, 64)[:n]
        f32, ni, nv = orc.compute_forces(*pos, ch, nm, com, lo, hi, theta, 1.0, soft, 1.0, n,
                                         precision=31 if comp else 32, mac=1,
                                         theta_g=per_particle, counts=True, bucket=bucket,
                                         straddle_mass=mu)
        g = host(gs)
        assert int(g[0]) == int(ni.sum()) and int(g[1]) == int(nv.sum())
        per = g[4:].reshape(W, 3)
        pad = np.zeros(W * 64 - n, dtype=np.int64)
        assert np.array_equal(per[:, 0], np.concatenate([ni, pad]).reshape(W, 64).sum(1))
        assert np.array_equal(per[:, 1], np.concatenate([nv, pad]).reshape(W, 64).sum(1))
        assert rel_rms(host(force), f32) <= 2e-6


def test_traverse_straddle_group(orc, torch_mod, ics32):
    """Per-group straddle masses (AABB test, 16-leaf buckets): each group's
    interactions and visits equal the oracle's with that group's straddle
    mass for its particles, forces to FP32 rounding; constant per-group
    values give exactly the scalar limit's results."""
    torch = torch_mod
    from src.kernels.physics_kernels import traverse
    from src.kernels.tree_kernels import build_tree_spans, normalize_pack_buckets
    pos, mass, ch, nm, com, lo, hi = _case(orc, ics32)
    _, codes = sorted_case(orc, ics32)
    n = len(mass)
    m = 2 * n - 1
    children = torch.full((m, 2), -1, dtype=torch.int32, device="cuda")
    parents = torch.full((m,), -1, dtype=torch.int32, device="cuda")
    far = torch.zeros(n - 1, dtype=torch.int32, device="cuda")
    build_tree_spans(dev(codes, torch), children, parents, far)
    _, csum, _, _ = orc.compute_multipoles(*pos, mass, ch, orc.build_tree(codes)[1])
    nodes = normalize_pack_buckets(children, far, dev(nm, torch), dev(csum, torch),
                                   dev(lo, torch), dev(hi, torch), 1.0, 16)
    root = torch.tensor([n], dtype=torch.int32, device="cuda")
    P = [dev(p, torch) for p in pos]
    W = (n + 63) // 64
    soft = 0.02 / 32
    sg = np.random.default_rng(8).choice([0.0, 2e-4, 1e-3, 4e-3], W).astype(np.float32)
    for theta, group_mu in ((0.5, sg), (0.7, np.full(W, 1e-3, np.float32))):
        force = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
        gs = torch.zeros(4 + 3 * W, dtype=torch.int64, device="cuda")
        traverse(*P, nodes, root, force, theta, 1.0, soft, 1.0, mac=1, stats=gs,
                 group_stats=True, straddle_group=dev(group_mu, torch))
        f32, ni, nv = orc.compute_forces(*pos, ch, nm, com, lo, hi, theta, 1.0, soft, 1.0, n,
                                         precision=32, mac=1, counts=True, bucket=16,
                                         straddle_g=np.repeat(group_mu, 64)[:n])
        g = host(gs)
        per = g[4:].reshape(W, 3)
        pad = np.zeros(W * 64 - n, dtype=np.int64)
        assert np.array_equal(per[:, 0], np.concatenate([ni, pad]).reshape(W, 64).sum(1))
        assert np.array_equal(per[:, 1], np.concatenate([nv, pad]).reshape(W, 64).sum(1))
        assert rel_rms(host(force), f32) <= 2e-6
        if np.all(group_mu == group_mu[0]):
            f2 = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
            traverse(*P, nodes, root, f2, theta, 1.0, soft, 1.0, mac=1,
                     straddle_mass=float(group_mu[0]))
            assert torch.equal(force, f2)


@pytest.mark.parametrize("mac", [0, 1])
def test_traverse_pairs_identical(orc, torch_mod, ics32, mac):
    """Two particles per lane (128-particle groups) gives the same forces bit
    for bit and the same per-particle interaction and visit counts."""
    torch = torch_mod
    from src.kernels.physics_kernels import traverse
    from src.kernels.tree_kernels import pack_nodes
    pos, mass, ch, nm, com, lo, hi = _case(orc, ics32)
    n = len(mass)
    nodes = pack_nodes(dev(ch, torch), dev(nm, torch), dev(com, torch), dev(lo, torch),
                       dev(hi, torch), 1.0, mac)
    root = torch.tensor([n], dtype=torch.int32, device="cuda")
    P = [dev(p, torch) for p in pos]
    tg = torch.from_numpy(np.random.default_rng(4).uniform(0.3, 0.8, (n + 63) // 64)
                          .astype(np.float32)).cuda()
    for group_theta in (None, tg):
        res = []
        for pairs in (False, True):
            f = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
            st = torch.zeros(4, dtype=torch.int64, device="cuda")
            # an odd-sized slice exercises partial groups and half-empty waves
            traverse(*P, nodes, root, f, 0.5, 1.0, 0.02 / 32, 1.0, off=37, count=n - 37 - 70,
                     mac=mac, theta_group=None if group_theta is None else group_theta,
                     stats=st, pairs=pairs)
            res.append((host(f), host(st)))
        assert np.array_equal(res[0][0], res[1][0])
        assert np.array_equal(res[0][1][:2], res[1][1][:2])
        assert res[1][1][2] < res[0][1][2]  # fewer wavefront pops with pairs


def test_traverse_launch_order(orc, torch_mod, ics32):
    """Launching the groups in theta order changes nothing in the results:
    forces bit for bit, per-group counts; order_groups is a permutation with
    non-decreasing theta bins."""
    torch = torch_mod
    from src.kernels.physics_kernels import order_groups, traverse
    from src.kernels.tree_kernels import pack_nodes
    pos, mass, ch, nm, com, lo, hi = _case(orc, ics32)
    n = len(mass)
    W = (n + 63) // 64
    nodes = pack_nodes(dev(ch, torch), dev(nm, torch), dev(com, torch), dev(lo, torch),
                       dev(hi, torch), 1.0, 1)
    root = torch.tensor([n], dtype=torch.int32, device="cuda")
    P = [dev(p, torch) for p in pos]
    tg = torch.from_numpy(np.random.default_rng(5).choice([0.1, 0.2, 0.3, 0.5, 0.8], W)
                          .astype(np.float32)).cuda()
    order = order_groups(tg)
    o = host(order)
    assert np.array_equal(np.sort(o), np.arange(W))
    assert np.all(np.diff(host(tg)[o]) >= 0)
    res = []
    for go in (None, order):
        f = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
        st = torch.zeros(4 + 3 * W, dtype=torch.int64, device="cuda")
        traverse(*P, nodes, root, f, 0.5, 1.0, 0.02 / 32, 1.0, mac=1, theta_group=tg, stats=st,
                 group_stats=True, group_order=go)
        res.append((host(f), host(st)))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])


def test_traverse_group_pops_and_longest_first(orc, torch_mod, ics32):
    """The AABB walk stores each group's stack pops (equal to the counting
    walk's per-group pops); order_groups_by_pops gives a permutation with
    non-increasing pop bins, and launching in it changes no result bit.
    The reference-test walk refuses group_pops.  An engine with
    LONGEST_FIRST ends three steps bit for bit where one without does."""
    torch = torch_mod
    from src.kernels.physics_kernels import order_groups_by_pops, traverse
    from src.kernels.tree_kernels import pack_nodes
    pos, mass, ch, nm, com, lo, hi = _case(orc, ics32)
    n = len(mass)
    W = (n + 63) // 64
    nodes = pack_nodes(dev(ch, torch), dev(nm, torch), dev(com, torch), dev(lo, torch),
                       dev(hi, torch), 1.0, 1)
    root = torch.tensor([n], dtype=torch.int32, device="cuda")
    P = [dev(p, torch) for p in pos]
    pops = torch.zeros(W, dtype=torch.int32, device="cuda")
    gs = torch.zeros(4 + 3 * W, dtype=torch.int64, device="cuda")
    f0 = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    traverse(*P, nodes, root, f0, 0.4, 1.0, 0.02 / 32, 1.0, mac=1, stats=gs, group_stats=True,
             group_pops=pops, straddle_mass=1e-4)
    per = host(gs)[4:].reshape(W, 3)
    assert np.array_equal(host(pops).astype(np.int64), per[:, 2])
    order = order_groups_by_pops(pops)
    o = host(order)
    assert np.array_equal(np.sort(o), np.arange(W))
    # heaviest first, to within one bin (a third of an octave): no later
    # group exceeds an earlier one by more than that
    po = 1.0 + host(pops)[o].astype(np.float64)
    later_max = np.maximum.accumulate(po[::-1])[::-1]
    assert np.all(later_max <= po * 2.0 ** (1.0 / 3.0) * (1.0 + 1e-5))
    # scaled by the change of each group's candidate cost: the estimate
    # pops * cost[k_now] / cost[k_prev] orders the groups (to within a bin);
    # candidate -1 (none) keeps the plain pops
    rng = np.random.default_rng(11)
    cost = torch.tensor([8.0, 2.0, 1.0, 0.5], dtype=torch.float32, device="cuda")
    kn = rng.integers(-1, 4, W).astype(np.int32)
    kp = rng.integers(-1, 4, W).astype(np.int32)
    so = host(order_groups_by_pops(pops, None, dev(kn, torch), dev(kp, torch), cost))
    assert np.array_equal(np.sort(so), np.arange(W))
    c = host(cost).astype(np.float64)
    scale = np.where((kn >= 0) & (kp >= 0), c[np.maximum(kn, 0)] / c[np.maximum(kp, 0)], 1.0)
    est = 1.0 + host(pops).astype(np.float64)[so] * scale[so]
    later_max = np.maximum.accumulate(est[::-1])[::-1]
    assert np.all(later_max <= est * 2.0 ** (1.0 / 3.0) * (1.0 + 1e-5))
    f1 = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    pops2 = torch.zeros(W, dtype=torch.int32, device="cuda")
    traverse(*P, nodes, root, f1, 0.4, 1.0, 0.02 / 32, 1.0, mac=1, group_order=order,
             group_pops=pops2, straddle_mass=1e-4)
    assert np.array_equal(host(f0), host(f1))
    assert np.array_equal(host(pops), host(pops2))
    ref_nodes = pack_nodes(dev(ch, torch), dev(nm, torch), dev(com, torch), dev(lo, torch),
                           dev(hi, torch), 1.0, 0)
    with pytest.raises(RuntimeError):
        traverse(*P, ref_nodes, root, f1, 0.4, 1.0, 0.02 / 32, 1.0, mac=0, group_pops=pops)
    # the engine
    from src.core.engine import SimulationEngine
    from src.core.state import SimulationState
    from src.utils import config as cfg
    runs = []
    for lf in (False, True):
        c = cfg.preset("small_32", MAC="aabb", BUCKET=16, STRADDLE_MASS=1e-4, THETA=0.4,
                       LONGEST_FIRST=lf)
        p0, v0 = zeldovich(32, c.SEED)
        eng = SimulationEngine(c, SimulationState(len(p0[0]), p0, v0))
        assert eng.longest_first == lf
        eng.run(3)
        runs.append([host(t) for t in eng.state.pos + eng.state.vel])
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


def test_traverse_group_theta_and_compensated(orc, torch_mod, ics32):
    torch = torch_mod
    from src.kernels.physics_kernels import traverse
    from src.kernels.tree_kernels import pack_nodes
    pos, mass, ch, nm, com, lo, hi = _case(orc, ics32)
    n = len(mass)
    soft = 0.02 / 32
    g = (n + 63) // 64
    tg = np.random.default_rng(3).uniform(0.3, 0.9, g).astype(np.float32)
    nodes = pack_nodes(dev(ch, torch), dev(nm, torch), dev(com, torch), dev(lo, torch),
                       dev(hi, torch), 1.0, 1)
    root = torch.tensor([n], dtype=torch.int32, device="cuda")
    P = [dev(p, torch) for p in pos]
    out = {}
    for comp in (False, True):
        force = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
        stats = torch.zeros(4, dtype=torch.int64, device="cuda")
        traverse(*P, nodes, root, force, 0.5, 1.0, soft, 1.0, mac=1, theta_group=dev(tg, torch),
                 compensated=comp, stats=stats)
        out[comp] = (host(force), host(stats))
    per_particle = np.repeat(tg, 64)[:n]
    f32, ni, nv = orc.compute_forces(*pos, ch, nm, com, lo, hi, 0.5, 1.0, soft, 1.0, n,
                                     precision=32, mac=1, theta_g=per_particle, counts=True)
    f64 = orc.compute_forces(*pos, ch, nm, com, lo, hi, 0.5, 1.0, soft, 1.0, n, precision=64,
                             mac=1, theta_g=per_particle)
    for comp in (False, True):
        assert int(out[comp][1][0]) == int(ni.sum())
    assert rel_rms(out[False][0], f32) <= 2e-6
    # compensation can only help against the FP64 arithmetic
    assert rel_rms(out[True][0], f64) <= rel_rms(out[False][0], f64) * 1.05


def test_direct_sum_matches_cpu(orc, torch_mod, ics16):
    torch = torch_mod
    from src.kernels.physics_kernels import direct_sum
    pos = ics16
    n = len(pos[0])
    mass = np.full(n, 1.0 / n, np.float32)
    tg = np.arange(0, n, 7)
    g = host(direct_sum(*[dev(p, torch) for p in pos], dev(mass, torch), dev(tg, torch), 1.0,
                        0.02 / 16))
    c = orc.direct_sum(*pos, mass, tg, 1.0, 0.02 / 16)
    assert rel_rms(g, c) < 1e-12


# --------------------------------------------------------- update / render

def test_integrate_bit_exact(orc, torch_mod, ics32):
    torch = torch_mod
    from src.kernels.physics_kernels import integrate
    pos, vel = zeldovich(32)
    n = len(pos[0])
    force = np.random.default_rng(9).standard_normal((n, 3)).astype(np.float32)
    a, H, dt = 0.02, 23970.3, 6.54e-7
    P = [dev(p, torch) for p in pos]
    V = [dev(v, torch) for v in vel]
    integrate[(n + 255) // 256, 256](*P, *V, dev(force, torch), a, H, dt)
    ref = orc.integrate(*pos, *vel, force, a, H, dt)
    for t, r in zip(P + V, ref):
        assert np.array_equal(host(t), r)


@pytest.mark.parametrize("res", [256, 1024])
def test_render_bit_exact(orc, torch_mod, ics32, res):
    torch = torch_mod
    from src.kernels.render_kernels import render_density
    from src.utils import config as cfg
    from src.utils.visualizer import get_mvp_matrix
    c = cfg.preset("repo_128")
    for step in (0, 777):
        mvp = get_mvp_matrix(step, (0.5, 0.5, 0.5), 1.0, c)
        grid = torch.zeros((res, res), dtype=torch.int64, device="cuda")
        n = len(ics32[0])
        render_density[(n + 255) // 256, 256](*[dev(p, torch) for p in ics32], grid, mvp, res)
        ref = orc.render_density(*ics32, mvp, res)
        assert np.array_equal(host(grid, np.uint64), ref)


# --------------------------------------------------------------- end to end

def _run_engine(preset, n_grid, steps, seed=42, **kw):
    from src.core.engine import SimulationEngine
    from src.core.state import SimulationState
    from src.utils import config as cfg
    c = cfg.preset(preset, N_GRID=n_grid, SOFTENING=0.02 / n_grid, SEED=seed, **kw)
    pos, vel = zeldovich(n_grid, seed)
    st = SimulationState(len(pos[0]), pos, vel)
    eng = SimulationEngine(c, st)
    for _ in range(steps):
        eng.step()
        eng.camera.capture(eng.steps_done - 1, *st.pos)
    return eng, c, pos, vel


@pytest.mark.parametrize("res", [64, 1024])
def test_colorize_frame(orc, torch_mod, ics32, res):
    """Fused colour mapping against the numpy float32 restatement of
    visualizer.py:115-171 on the splat of a real frame.  log1p, the division
    by the maximum and pow each differ by an ulp or two between libraries, so
    val = img^2.2 carries up to about 8 ulps (<= 1e-6 absolute, val <= 1);
    the steepest palette segment multiplies that by 1.3 / 0.1 = 13, hence
    1.5e-5 on the floats.  Bytes are equal except where a value sits at a
    truncation boundary."""
    torch = torch_mod
    from src.kernels.render_kernels import render_density, to_float
    from src.utils import config as cfg
    from src.utils.visualizer import colorize_frame, get_mvp_matrix
    c = cfg.preset("small_32", RES=res)
    grid = torch.zeros((res, res), dtype=torch.int64, device="cuda")
    mvp = get_mvp_matrix(3, (0.5, 0.5, 0.5), 1.0, c)
    render_density(*[dev(p, torch) for p in ics32], grid, mvp, res)
    frame = colorize_frame(grid)
    rgb = torch.zeros((res, res, 3), dtype=torch.uint8, device="cuda")
    from src.kernels.render_kernels import colorize
    colorize(grid, rgb=rgb)
    ref_f, ref_b = orc.colorize_frame(to_float(grid).float().cpu().numpy())
    f = host(frame)
    assert np.abs(f - ref_f).max() <= 1.5e-5
    b = host(rgb).astype(np.int32)
    diff = np.abs(b - ref_b.astype(np.int32))
    assert diff.max() <= 1 and np.mean(diff > 0) < 1e-3
    assert np.array_equal(b, (f * np.float32(255)).astype(np.uint8))
    # an empty grid is black (the reference's + 1e-9 guard)
    grid.zero_()
    assert host(colorize_frame(grid)).max() == 0.0


def test_engine_steps_match_oracle(orc, torch_mod):
    """C0: the five steps of small_32 (N = 32^3) on the GPU, each stage
    compared with the oracle on the same input (tests/step_check.py):
    codes, permutation, tree, FP64 moments, kick-drift and density grid bit
    for bit, accelerations of every particle within 1e-5 relative RMS of the
    reference method's FP64 arithmetic."""
    from src.core.engine import SimulationEngine
    from src.core.state import SimulationState
    from src.utils import config as cfg
    from step_check import check_steps
    c = cfg.preset("small_32")
    pos, vel = zeldovich(32, c.SEED)
    eng = SimulationEngine(c, SimulationState(len(pos[0]), pos, vel))
    errs = check_steps(orc, eng, c.RUN_STEPS)
    assert len(errs) == 5


def test_engine_deterministic(torch_mod):
    e1, *_ = _run_engine("small_32", 32, 3)
    e2, *_ = _run_engine("small_32", 32, 3)
    for a, b in zip(e1.state.pos + e1.state.vel, e2.state.pos + e2.state.vel):
        assert np.array_equal(host(a), host(b))
    assert np.array_equal(host(e1.camera.grid_d), host(e2.camera.grid_d))


@pytest.mark.parametrize("n", [1, 2, 3, 65])
def test_tiny_sizes_every_stage(orc, torch_mod, n):
    """Every stage at the smallest sizes, against the oracle on the same
    input: n = 1 (the root is the only leaf, id 0, so find_root returns 0
    and the walk visits it without an interaction), 2 and 3 (roots over
    leaves only, and over a leaf and an internal node), 65 (a full group
    and a one-particle group).  Codes, permutation, tree, root, FP64
    moments, kick-drift and density grid bit for bit; per-particle visits
    and interactions of both acceptance tests exact; forces of the walk and
    of the compute_forces drop-in (its split walk) within FP32 rounding of
    the FP32 oracle, and exactly zero for a lone particle."""
    torch = torch_mod
    from src.kernels.morton import compute_codes
    from src.kernels.physics_kernels import compute_forces, integrate, traverse
    from src.kernels.render_kernels import render_density
    from src.kernels.tree_kernels import find_root, normalize_com, pack_nodes
    from src.utils import config as cfg
    from src.utils.visualizer import get_mvp_matrix
    rng = np.random.default_rng(1000 + n)
    pos = [rng.uniform(0, 1, n).astype(np.float32) for _ in range(3)]
    vel = [rng.standard_normal(n).astype(np.float32) * np.float32(1e-3) for _ in range(3)]
    blocks = (n + 255) // 256
    ref_codes, ref_idx = orc.compute_codes(*pos)
    codes = torch.zeros(n, dtype=torch.int64, device="cuda")
    ind = torch.zeros(n, dtype=torch.int32, device="cuda")
    compute_codes[blocks, 256](*[dev(p, torch) for p in pos], codes, ind, n)
    assert np.array_equal(host(codes, np.uint64), ref_codes)
    assert np.array_equal(host(ind, np.uint32), ref_idx)
    sc, perm = _gpu_sort(torch, ref_codes)
    rperm = orc.argsort_stable(ref_codes)
    assert np.array_equal(perm, rperm)
    assert np.array_equal(sc, ref_codes[rperm])
    pos = [p[rperm] for p in pos]
    vel = [v[rperm] for v in vel]
    ch, par, cnt, root = _gpu_tree(torch, sc)
    ref_ch, ref_par = orc.build_tree(sc)
    assert np.array_equal(host(ch), ref_ch)
    assert np.array_equal(host(par), ref_par)
    M = 2 * n - 1
    find_root[(M + 255) // 256, 256](par, root)
    r = orc.find_root(ref_par)
    assert int(root.item()) == r == (0 if n == 1 else n)
    mass = np.full(n, 1.0 / n, np.float32)
    nm, nc, lo, hi = _gpu_moments(torch, pos, mass, ref_ch, ref_par)
    r_nm, r_nc, r_lo, r_hi = orc.compute_multipoles(*pos, mass, ref_ch, ref_par)
    for a, b in ((nm, r_nm), (nc, r_nc), (lo, r_lo), (hi, r_hi)):
        assert np.array_equal(host(a), b)
    normalize_com[(M + 255) // 256, 256](nm, nc, M)
    com = orc.normalize_com(r_nm, r_nc)
    assert np.array_equal(host(nc), com)
    soft = 1e-3
    P = [dev(p, torch) for p in pos]
    for mac in (0, 1):
        nodes = pack_nodes(ch, nm, nc, lo, hi, 1.0, mac)
        force = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
        stats = torch.zeros(4, dtype=torch.int64, device="cuda")
        traverse(*P, nodes, root, force, 0.5, 1.0, soft, 1.0, mac=mac, stats=stats)
        f32, ni, nv = orc.compute_forces(*pos, ref_ch, r_nm, com, r_lo, r_hi, 0.5, 1.0, soft,
                                         1.0, r, precision=32, mac=mac, counts=True)
        s = host(stats)
        assert int(s[0]) == int(ni.sum()) and int(s[1]) == int(nv.sum())
        if n == 1:
            assert int(ni[0]) == 0 and int(nv[0]) == 1
            assert not host(force).any()
        else:
            assert rel_rms(host(force), f32) <= 2e-6
    f32, _, _ = orc.compute_forces(*pos, ref_ch, r_nm, com, r_lo, r_hi, 0.5, 1.0, soft, 1.0, r,
                                   precision=32, counts=True)
    force = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    compute_forces[blocks, 256](*P, ch, nm, nc, lo, hi, force, 0.5, 1.0, soft, 1.0, root, None,
                                None, 0)
    if n == 1:
        assert not host(force).any()
    else:
        assert rel_rms(host(force), f32) <= 5e-6
    a, H, dt = 0.02, 23970.3, 6.54e-7
    V = [dev(v, torch) for v in vel]
    integrate[blocks, 256](*P, *V, force, a, H, dt)
    ref = orc.integrate(*pos, *vel, host(force), a, H, dt)
    for t, q in zip(P + V, ref):
        assert np.array_equal(host(t), q)
    mvp = get_mvp_matrix(3, (0.5, 0.5, 0.5), 1.0, cfg.preset("repo_128"))
    grid = torch.zeros((64, 64), dtype=torch.int64, device="cuda")
    render_density[blocks, 256](*P, grid, mvp, 64)
    assert np.array_equal(host(grid, np.uint64), orc.render_density(*ref[:3], mvp, 64))


def test_empty_inputs_are_no_ops(torch_mod):
    """n = 0 at every drop-in (with a one-block launch, as a caller with an
    empty batch would issue it): no error, nothing written."""
    torch = torch_mod
    from src.kernels.morton import compute_codes
    from src.kernels.physics_kernels import compute_forces, integrate
    from src.kernels.render_kernels import render_density
    from src.kernels.tree_kernels import build_tree, compute_multipoles
    from src.utils import config as cfg
    from src.utils.visualizer import get_mvp_matrix
    f0 = [torch.zeros(0, dtype=torch.float32, device="cuda") for _ in range(3)]
    codes = torch.zeros(0, dtype=torch.int64, device="cuda")
    ind = torch.zeros(0, dtype=torch.int32, device="cuda")
    compute_codes[1, 256](*f0, codes, ind, 0)
    ch = torch.zeros((0, 2), dtype=torch.int32, device="cuda")
    par = torch.zeros(0, dtype=torch.int32, device="cuda")
    build_tree[1, 256](codes, ch, par)
    nm = torch.zeros(0, dtype=torch.float64, device="cuda")
    nc = torch.zeros((0, 3), dtype=torch.float64, device="cuda")
    lo = torch.zeros((0, 3), dtype=torch.float32, device="cuda")
    hi = torch.zeros((0, 3), dtype=torch.float32, device="cuda")
    cnt = torch.zeros(0, dtype=torch.int32, device="cuda")
    compute_multipoles[1, 256](*f0, f0[0], ch, par, nm, nc, lo, hi, cnt)
    force = torch.zeros((0, 3), dtype=torch.float32, device="cuda")
    root = torch.zeros(1, dtype=torch.int32, device="cuda")
    compute_forces[1, 256](*f0, ch, nm, nc, lo, hi, force, 0.5, 1.0, 1e-3, 1.0, root, None, None,
                           0)
    V = [torch.zeros(0, dtype=torch.float32, device="cuda") for _ in range(3)]
    integrate[1, 256](*f0, *V, force, 0.02, 23970.3, 6.54e-7)
    grid = torch.zeros((64, 64), dtype=torch.int64, device="cuda")
    mvp = get_mvp_matrix(0, (0.5, 0.5, 0.5), 1.0, cfg.preset("repo_128"))
    render_density[1, 256](*f0, grid, mvp, 64)
    torch.cuda.synchronize()
    assert int(grid.abs().sum().item()) == 0
    assert int(root.item()) == 0


def test_full_size_properties(torch_mod):
    """repo_128 (N = 128^3): sorted codes, a permutation, a valid tree with
    root n and root mass 1, finite forces and a non-empty density grid."""
    torch = torch_mod
    eng, c, pos, vel = _run_engine("repo_128", 128, 1)
    st = eng.state
    n = st.n
    codes = host(st.codes, np.uint64)
    assert np.all(codes[1:] >= codes[:-1])
    perm = host(st.sort_idx, np.uint32)
    assert np.array_equal(np.sort(perm), np.arange(n, dtype=np.uint32))
    tm = eng.tree_manager
    assert tm.validate() == 0
    assert int(tm.root_idx.item()) == n
    assert abs(float(tm.mass[n].item()) - 1.0) < 1e-12
    f = host(st.force)
    assert np.all(np.isfinite(f))
    grid = host(eng.camera.grid_d, np.uint64)
    assert grid.sum() > 0
