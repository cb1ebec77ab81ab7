"""Stage-by-stage comparison of engine steps on the GPU with the oracle.

Each step starts from the GPU's own state (positions, velocities, masses),
so every stage is compared on identical inputs: Morton codes, the sort
permutation and the permuted particles, children, parents and root, the
FP64 node mass, the normalised FP64 CoM and the FP32 AABBs bit for bit;
accelerations within 1e-5 relative RMS of the oracle's FP64 arithmetic of
the reference method (all particles, or sampled targets); the kick-drift
update bit for bit given the GPU's accelerations; and the fixed-point
density grid bit for bit.
"""
import numpy as np

from conftest import rel_rms


def host(t, dtype=None):
    a = t.detach().cpu().numpy()
    return a.view(dtype) if dtype is not None else a


def check_steps(orc, eng, steps, targets=None, log=print):
    """Runs `steps` engine steps with checks; returns the per-step force errors."""
    from src.kernels.physics_kernels import integrate
    st, c, tm = eng.state, eng.config, eng.tree_manager
    n = st.n
    errs = []
    for k in range(steps):
        pos = [host(t).copy() for t in st.pos]
        vel = [host(t).copy() for t in st.vel]
        mass = host(st.mass).copy()
        H, dt = eng.timestep()
        st.sort_buffer_by_morton(c)
        codes, _ = orc.compute_codes(*pos)
        perm = orc.argsort_stable(codes)
        assert np.array_equal(host(st.sort_idx, np.uint32), perm), f"step {k}: permutation"
        assert np.array_equal(host(st.codes, np.uint64), codes[perm]), f"step {k}: codes"
        p = [x[perm] for x in pos]
        v = [x[perm] for x in vel]
        m = mass[perm]
        for a, b in zip(st.pos + st.vel + (st.mass,), p + v + [m]):
            assert np.array_equal(host(a), b), f"step {k}: permuted particles"
        tm.build_packed(st.pos, st.mass, st.codes, c.BOX_SIZE, eng.mac, eng.bucket)
        ch, par = orc.build_tree(codes[perm])
        assert np.array_equal(host(tm.child), ch), f"step {k}: children"
        assert np.array_equal(host(tm.parent), par), f"step {k}: parents"
        assert int(tm.root_idx.item()) == orc.find_root(par) == n
        nm, nc, lo, hi = orc.compute_multipoles(*p, m, ch, par)
        com = orc.normalize_com(nm, nc)
        assert np.array_equal(host(tm.mass), nm), f"step {k}: node mass"
        assert np.array_equal(host(tm.com), com), f"step {k}: node CoM"
        assert np.array_equal(host(tm.aabb_min), lo), f"step {k}: node min"
        assert np.array_equal(host(tm.aabb_max), hi), f"step {k}: node max"
        eng.compute_forces()
        f = host(st.force)
        assert np.all(np.isfinite(f))
        if targets is None:
            ref = orc.compute_forces(*p, ch, nm, com, lo, hi, c.THETA, c.G, c.SOFTENING,
                                     c.BOX_SIZE, n, precision=64)
            e = rel_rms(f, ref)
        else:
            ref = orc.compute_forces(*p, ch, nm, com, lo, hi, c.THETA, c.G, c.SOFTENING,
                                     c.BOX_SIZE, n, precision=64, targets=targets)
            e = rel_rms(f[targets], ref)
        errs.append(e)
        assert e <= 1e-5, f"step {k}: force relative RMS {e:.3e}"
        integrate(*st.pos, *st.vel, st.force, eng.a, H, dt)
        upd = orc.integrate(*p, *v, f, eng.a, H, dt)
        for a, b in zip(st.pos + st.vel, upd):
            assert np.array_equal(host(a), b), f"step {k}: kick-drift"
        grid = eng.camera.capture(eng.steps_done, *st.pos)
        from src.utils.visualizer import get_mvp_matrix
        mvp = get_mvp_matrix(eng.steps_done, eng.camera.box_center, c.BOX_SIZE, c)
        ref_grid = orc.render_density(*upd[:3], mvp, eng.camera.res)
        assert np.array_equal(host(grid, np.uint64), ref_grid), f"step {k}: density grid"
        eng.a += eng.da
        eng.steps_done += 1
        log(f"step {k}: bit-exact stages, force relative RMS {e:.2e}")
    return errs
