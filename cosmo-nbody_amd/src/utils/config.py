"""Simulation configuration.

Module-level constants carry the reference's defaults with the same names
(src/utils/config.py:9-35), so ``import src.utils.config as consts`` and
``SimulationEngine(consts, state)`` work as in the reference.  ``Config`` is
a dataclass with the same attribute names plus the knobs this build adds;
``preset()`` returns the named workloads:

  reference_defaults  the reference's own constants (theta 0.7, eps 1e-3)
  small_32            N = 32^3 Zel'dovich ICs, theta 0.5, eps 0.02/32, 5 steps
  repo_128            N = 128^3, theta 0.5, eps 0.02/128, 20 steps, 1024^2 splat
  clustered_128       N = 128^3 evolved 300 steps, theta swept (precision study)
  repo_256            N = 256^3 Zel'dovich ICs, theta 0.5, eps 0.02/256, 50 steps,
                      2048^2 splat (the reference test at the large size)
  halos_128           N = 128^3 heterogeneous test state (half in 64 Plummer
                      spheres, half uniform), AABB acceptance test with
                      32-leaf buckets, compensated sums, straddle mass 5e-5
  learned_128         halos_128 with the learned per-group opening angle and
                      straddle mass (network THETA_NETWORK)
  large_256           N = 256^3 halo state, 50 steps, 2048^2 splat, the learned
                      selection (the same test, straddle mass 6.25e-6: the same
                      105 particle masses as at 128^3)
  halos_256           the same as large_256 (the name the learned-theta
                      evaluation uses)

The learned selection is trained and calibrated on the halo state, the
strongly clustered state of these workloads (the Zel'dovich ICs stay nearly
uniform under the reference schedule); on a near-uniform state the same
network selects large opening angles whose error is not calibrated.

Time stepping follows the reference schedule: da = (A_END - A_START) /
N_STEPS with N_STEPS = 2500 (src/core/engine.py:46); RUN_STEPS is how many of
those steps a workload runs.
"""
from __future__ import annotations

import dataclasses

# --- SIMULATION PARAMETERS (reference src/utils/config.py:9-14) ---
N_GRID = 128
N_BODIES = N_GRID ** 3
BOX_SIZE = 1.0
SOFTENING = 0.001
N_STEPS = 2500
A_START, A_END = 0.02, 1.0

# --- RENDER SETTINGS (:16-21) ---
ROTATION_SPEED = 0.001
CAM_DIST_MULT = 1.8
FOV = 60.0
RES = 1024
OUTPUT_DIR = "render_output"

# --- LAUNCH CONFIG (:23-27); TPB/BPG only size the [off, off+B*T) range ---
TPB = 256
BPG = (N_BODIES + TPB - 1) // TPB
BATCH_SIZE = int(5_000)
CACHE_SIZE = 256

# --- PHYSICS CONSTANTS (:29-35) ---
THETA = 0.7
G = 1.0
OMEGA_M = 1.0
OMEGA_L = 0.692
H0 = 67.8
STACK_SIZE_CHILD = 128

# --- additions of this build ---
SEED = 42                  # Zel'dovich phase seed (reference cosmology.py:25)
MORTON_MASKS = "full21"    # "full21" (21 bits/axis) or "ref10" (morton.py:33-36 as written)
RUN_STEPS = N_STEPS
RENDER_FRAC_BITS = 20      # fixed-point fraction bits of the density bins
COMPENSATED = False        # Neumaier-compensated far-field accumulation
MAC = "ref"                # acceptance test: "ref" (physics_kernels.py:128-133) or "aabb"
BUCKET = 0                 # AABB test: nodes of <= BUCKET leaves are summed leaf by leaf when opened
STRADDLE_MASS = 0.0        # AABB test: nodes lighter than STRADDLE_MASS pass the
                           # minimum-image guard (0 = exact guard)
TRAVERSAL_PAIRS = False    # two particles per lane in the traversal (same results, slower at 128^3)
LONGEST_FIRST = False      # AABB test: launch the traversal's groups by decreasing stack
                           # pops of the previous launch (pays on clustered states; same results)
LEARNED_THETA = False
THETA_NETWORK = ""         # trained per-group theta network (LearnedTheta.save file, a path
                           # or a name in cosmo-nbody_amd/networks/) for LEARNED_THETA
ZA_AMPLITUDE = 0.3         # rms Zel'dovich displacement in grid spacings (cosmology.py:49)
IC_KIND = "zeldovich"      # "zeldovich" (reference) or "halos" (heterogeneous test state)
HALO_FRACTION = 0.5
N_HALOS = 64
HALO_RADII = (0.003, 0.03)


@dataclasses.dataclass
class Config:
    N_GRID: int = N_GRID
    BOX_SIZE: float = BOX_SIZE
    SOFTENING: float = SOFTENING
    N_STEPS: int = N_STEPS
    A_START: float = A_START
    A_END: float = A_END
    ROTATION_SPEED: float = ROTATION_SPEED
    CAM_DIST_MULT: float = CAM_DIST_MULT
    FOV: float = FOV
    RES: int = RES
    OUTPUT_DIR: str = OUTPUT_DIR
    TPB: int = TPB
    BATCH_SIZE: int = BATCH_SIZE
    CACHE_SIZE: int = CACHE_SIZE
    THETA: float = THETA
    G: float = G
    OMEGA_M: float = OMEGA_M
    OMEGA_L: float = OMEGA_L
    H0: float = H0
    STACK_SIZE_CHILD: int = STACK_SIZE_CHILD
    SEED: int = SEED
    MORTON_MASKS: str = MORTON_MASKS
    RUN_STEPS: int = RUN_STEPS
    ZA_AMPLITUDE: float = ZA_AMPLITUDE
    IC_KIND: str = IC_KIND
    HALO_FRACTION: float = HALO_FRACTION
    N_HALOS: int = N_HALOS
    HALO_RADII: tuple = HALO_RADII
    RENDER_FRAC_BITS: int = RENDER_FRAC_BITS
    COMPENSATED: bool = COMPENSATED
    MAC: str = MAC
    BUCKET: int = BUCKET
    STRADDLE_MASS: float = STRADDLE_MASS
    TRAVERSAL_PAIRS: bool = TRAVERSAL_PAIRS
    LONGEST_FIRST: bool = LONGEST_FIRST
    LEARNED_THETA: bool = LEARNED_THETA
    THETA_NETWORK: str = THETA_NETWORK
    NAME: str = "reference_defaults"

    @property
    def N_BODIES(self) -> int:
        return self.N_GRID ** 3

    @property
    def BPG(self) -> int:
        return (self.N_BODIES + self.TPB - 1) // self.TPB

    @property
    def DA(self) -> float:
        return (self.A_END - self.A_START) / self.N_STEPS

    def replace(self, **kw) -> "Config":
        return dataclasses.replace(self, **kw)


def _plain(n_grid: int, **kw) -> Config:
    base = dict(N_GRID=n_grid, THETA=0.5, SOFTENING=0.02 / n_grid)
    base.update(kw)
    return Config(**base)


PRESETS = {
    "reference_defaults": lambda: Config(NAME="reference_defaults"),
    "small_32": lambda: _plain(32, RUN_STEPS=5, NAME="small_32"),
    "repo_128": lambda: _plain(128, RUN_STEPS=20, RES=1024, NAME="repo_128"),
    "clustered_128": lambda: _plain(128, RUN_STEPS=300, NAME="clustered_128"),
    "repo_256": lambda: _plain(256, RUN_STEPS=50, RES=2048, NAME="repo_256"),
    "learned_128": lambda: _plain(128, RUN_STEPS=20, LEARNED_THETA=True, MAC="aabb", BUCKET=32,
                                  IC_KIND="halos", COMPENSATED=True, STRADDLE_MASS=5e-5,
                                  THETA_NETWORK="halos_128.npz", LONGEST_FIRST=True,
                                  NAME="learned_128"),
    "large_256": lambda: _plain(256, RUN_STEPS=50, RES=2048, LEARNED_THETA=True, MAC="aabb",
                                IC_KIND="halos", BUCKET=32, COMPENSATED=True,
                                STRADDLE_MASS=6.25e-6, THETA_NETWORK="halos_128.npz",
                                LONGEST_FIRST=True,
                                NAME="large_256"),
    "halos_128": lambda: _plain(128, RUN_STEPS=20, MAC="aabb", IC_KIND="halos", BUCKET=32,
                                COMPENSATED=True, STRADDLE_MASS=5e-5, LONGEST_FIRST=True,
                                NAME="halos_128"),
    "halos_256": lambda: _plain(256, RUN_STEPS=50, RES=2048, MAC="aabb", IC_KIND="halos",
                                BUCKET=32, COMPENSATED=True, STRADDLE_MASS=6.25e-6,
                                LEARNED_THETA=True, THETA_NETWORK="halos_128.npz",
                                LONGEST_FIRST=True, NAME="halos_256"),
}


def preset(name: str, **overrides) -> Config:
    if name not in PRESETS:
        raise KeyError(f"unknown preset {name!r}; known: {sorted(PRESETS)}")
    return PRESETS[name]().replace(**overrides)
