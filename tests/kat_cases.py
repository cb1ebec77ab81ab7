"""Hand-built known-answer cases for the acceptance test and the splat
orientation of the reference (shared by the oracle tests and the GPU tests).

Each force case is a three-particle tree written out by hand: leaf 0 is the
target T, leaves 1 and 2 are A and B, internal node 3 (the root) has
children (T, node 4) and node 4 has children (A, B).  The expected visits,
interactions and acceleration of T are derived by hand from the text of
src/kernels/physics_kernels.py:115-149:

  sx = max_x - min_x; sx = box if sx > box/2; accept a node when it is a
  leaf or sx^2 < theta^2 d^2 (d to the CoM, minimum image); interact when
  d^2 > 1e-7 with G m (d^2 + eps^2)^-3/2 d; otherwise push c2 then c1.

Node masses and CoMs are what the reference stores: FP32 roundings of the
FP64 sums (src/core/octree.py:37-40).
"""
import numpy as np

SOFT = 1e-3
G = 1.0


def tree_arrays(points, masses):
    """(pos SoA float32, mass float32, children int32[5,2], parents int32[5])
    of the fixed topology root 3 = (0, 4), node 4 = (1, 2)."""
    P = np.asarray(points, np.float32)
    pos = [np.ascontiguousarray(P[:, k]) for k in range(3)]
    mass = np.asarray(masses, np.float32)
    children = np.array([[-1, -1], [-1, -1], [-1, -1], [0, 4], [1, 2]], np.int32)
    parents = np.array([3, 4, 4, -1, 3], np.int32)
    return pos, mass, children, parents


def _term(m, d):
    d = np.asarray(d, np.float64)
    return G * float(m) * d / (d @ d + SOFT * SOFT) ** 1.5


def _min_image(d):
    d = np.asarray(d, np.float64)
    return d - np.rint(d)


def _pair_node(points, masses):
    """FP32 mass and CoM of node 4 = (A, B) as the reference stores them."""
    P = np.asarray(points, np.float32).astype(np.float64)
    m = np.asarray(masses, np.float32).astype(np.float64)
    M = m[1] + m[2]
    com = (m[1] * P[1] + m[2] * P[2]) / M
    return float(np.float32(M)), np.float32(com).astype(np.float64)


# Thin in x, tall in y.  Node 4 = (A, B) has x-extent 0 and y-extent 0.1 at
# distance 0.4 from T.  At theta = 0.2: x-extent test 0 < 0.04 * 0.16 ->
# accepted (one interaction with the pair's monopole); a largest-extent test
# (0.1^2 = 0.01 > 0.0064) would open it and interact with A and B.  The root
# (x-extent 0.4, CoM 0.2667 from T) is opened: 0.16 > 0.04 * 0.0711.
XEXT = dict(points=[(0.2, 0.5, 0.5), (0.6, 0.45, 0.5), (0.6, 0.55, 0.5)],
            masses=[1 / 3, 1 / 3, 1 / 3], theta=0.2, visits=3, interactions=1)

# x-extent above box/2.  Node 4 = (A, B) spans x from 0.05 to 0.6 (0.55 >
# 0.5), so sx is replaced by box = 1 and 1 < theta^2 d^2 is impossible for
# theta <= 1 (d^2 <= 3/4 under the minimum image): T opens it and interacts
# with A and B.  Without the clamp, 0.55^2 = 0.3025 < 1 * 0.4329 would accept
# it (d = (0, -0.45, -0.48) to its CoM).  The root spans 0.55 too (opened).
CLAMP = dict(points=[(0.325, 0.55, 0.58), (0.05, 0.1, 0.1), (0.6, 0.1, 0.1)],
             masses=[1 / 3, 1 / 3, 1 / 3], theta=1.0, visits=5, interactions=2)

# Just below box/2: node 4 spans x from 0.1 to 0.55 (0.45 <= 0.5, kept), and
# 0.2025 < 1 * 0.4329 accepts it at theta = 1.  The root (x-extent 0.45, CoM
# at d = (0, -0.3, -0.32), d^2 = 0.1924 < 0.2025) is opened.
NOCLAMP = dict(points=[(0.325, 0.55, 0.58), (0.1, 0.1, 0.1), (0.55, 0.1, 0.1)],
               masses=[1 / 3, 1 / 3, 1 / 3], theta=1.0, visits=3, interactions=1)

FORCE_CASES = {"x_extent": XEXT, "clamp": CLAMP, "no_clamp": NOCLAMP}


def expected_force(case):
    """Hand-derived acceleration of T (FP64) for the case."""
    P = np.asarray(case["points"], np.float32).astype(np.float64)
    m = np.asarray(case["masses"], np.float32)
    if case["interactions"] == 1:  # the monopole of node 4
        M, com = _pair_node(case["points"], case["masses"])
        return _term(M, _min_image(com - P[0]))
    return _term(m[1], _min_image(P[1] - P[0])) + _term(m[2], _min_image(P[2] - P[0]))


def opened_force(case):
    """What T would get if node 4 were opened (A and B separately)."""
    P = np.asarray(case["points"], np.float32).astype(np.float64)
    m = np.asarray(case["masses"], np.float32)
    return _term(m[1], _min_image(P[1] - P[0])) + _term(m[2], _min_image(P[2] - P[0]))


# ---------------------------------------------------------------- splat

def pinhole_pixel(point, res, step=0, box=1.0, dist_mult=1.8, fov=60.0):
    """(row, col, fractional parts) of a point seen by the reference's orbiting
    camera (src/utils/visualizer.py:81-112 and render_kernels.py:36-47),
    derived from the pinhole model in FP64: the eye at centre + (r cos a,
    0.2 box, r sin a) looks at the centre with up = +y; with z_cam = unit(eye
    - centre), x_cam = unit(up x z_cam) (screen right), y_cam = z_cam x
    x_cam (screen up), view coordinates v = (x_cam, y_cam, z_cam) . (p -
    eye), w = -v_z, f = 1/tan(fov/2): column = (f v_x / w + 1) res / 2 and
    row = (1 - f v_y / w) res / 2.  So a point right of the centre on
    screen has a larger column, and a point above it a smaller row."""
    c = np.full(3, box / 2, np.float64)
    a = step * 0.001
    r = box * dist_mult
    eye = c + np.array([r * np.cos(a), 0.2 * box, r * np.sin(a)])
    z = (eye - c) / np.linalg.norm(eye - c)
    x = np.cross([0.0, 1.0, 0.0], z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    v = np.asarray(point, np.float64) - eye
    vx, vy, vz = x @ v, y @ v, z @ v
    w = -vz
    f = 1.0 / np.tan(np.radians(fov) / 2)
    col = (f * vx / w + 1.0) * 0.5 * res
    row = (1.0 - f * vy / w) * 0.5 * res
    return int(np.floor(row)), int(np.floor(col)), row % 1.0, col % 1.0


# At step 0 the eye sits on the +x side of the box, so screen right is -z
# and screen up is +y: a point at larger z lies left of the centre (column
# < res/2, same row as the centre), a point at larger y lies above it (row
# < res/2, same column as the centre).
SPLAT_POINTS = {"left": (0.5, 0.5, 0.65), "up": (0.5, 0.65, 0.5),
                "right_down": (0.5, 0.42, 0.37)}
