"""Dam-break workloads of the benchmark configurations (geometry in metres).

* ``dambreak2d``: 2-D tank 1.61 x 0.8, column 0.4 wide x 0.3 high, dx 0.005
  (4800 fluid particles on a lattice + 3 dummy layers).
* ``dambreak3d``: 3-D tank 3.22 x 1.0 x 1.0, column 1.2 x 1.0 x 0.55, dx 0.008
  (1,293,750 fluid particles + 3 dummy layers on the floor and four walls).
  dx 0.016 gives the 2x-coarser 3-D case (~160k fluid particles).
* ``wavetank3d``: the floating-body configuration: tank 4.0 x 1.0 x 1.2 filled
  to 0.6, dx 0.006 (~11.1M fluid particles), a piston wavemaker at x = 0 with a
  0.05 m peak-to-peak stroke at 1 Hz (amplitude 0.025), and a 0.3 x 0.3 x 0.2
  box of density 500 (9 kg) at mid-tank, placed at its equilibrium draft 0.1.
  Coarser dx gives the same geometry with fewer particles.
* ``ensemble2d``: the batched ensemble, 512 independent 2-D dam breaks in a
  1.61 x 1.0 tank, scene k drawn from seed + k through SplitMix64: column
  width U(0.2, 0.6), height U(0.3, 0.8), dx U(0.004, 0.01) (600 to 30,000
  fluid particles per scene).
All use the Wendland C2 kernel with h = 1.3 dx, Tait EOS gamma 7 with
rho0 = 1000 and c0 = 10 sqrt(2 g H), delta 0.1, alpha 0.02, CFL 0.2.
"""
import numpy as np

from . import Batch, Case, DIFF_ANTUONO, rigid_state

_M64 = (1 << 64) - 1


def dambreak2d(dx=0.005, gamma=7.0, diffusion=DIFF_ANTUONO, lx=1.61, lz=0.8, cw=0.4, ch=0.3, layers=3, **kw):
    return Case.dambreak(2, dx, lx, 0.0, lz, cw, ch, layers=layers, gamma=gamma, cfl=0.2,
                         diffusion=diffusion, **kw)


def dambreak3d(dx=0.008, gamma=7.0, diffusion=DIFF_ANTUONO, lx=3.22, wy=1.0, lz=1.0, cw=1.2, ch=0.55, layers=3,
               **kw):
    return Case.dambreak(3, dx, lx, wy, lz, cw, ch, layers=layers, gamma=gamma, cfl=0.2, diffusion=diffusion,
                         **kw)


PISTON_AMPLITUDE, PISTON_FREQUENCY = 0.025, 1.0
BOX_DENSITY = 500.0
BOX = ((1.85, 0.35, 0.5), (2.15, 0.65, 0.7))


def wavetank3d(dx=0.006, gamma=7.0, diffusion=DIFF_ANTUONO, lx=4.0, wy=1.0, lz=1.2, depth=0.6, box=BOX, layers=3,
               piston_travel=0.05, **kw):
    """Wave tank with a piston and a floating box; set the body with box_body(*case.box, BOX_DENSITY)
    (the box as snapped to the lattice) and the piston with (PISTON_AMPLITUDE, PISTON_FREQUENCY)."""
    kw.setdefault("grid_size_factor", 2.0)  # cells of edge h: fewer candidates per neighbour walk
    return Case.wavetank(dx, lx, wy, lz, depth, box[0], box[1], layers=layers, piston_travel=piston_travel,
                         gamma=gamma, cfl=0.2, diffusion=diffusion, **kw)


def box_body(lo, hi, density, rho_c=None):
    """Rigid state of a solid box [lo, hi] of uniform `density` at rest: mass, centre and
    I = M/12 diag(b^2 + c^2, a^2 + c^2, a^2 + b^2)."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    a, b, c = hi - lo
    m = density * a * b * c
    inertia = np.diag([b * b + c * c, a * a + c * c, a * a + b * b]) * m / 12.0
    return rigid_state((lo + hi) / 2 if rho_c is None else rho_c, m, inertia)


def splitmix64(state):
    """One SplitMix64 step: (new state, 64-bit output)."""
    state = (state + 0x9E3779B97F4A7C15) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return state, z ^ (z >> 31)


def ensemble_draw(seed):
    """(column width, column height, dx) of one scene: three SplitMix64 uniforms from `seed`."""
    state, u = seed & _M64, []
    for _ in range(3):
        state, z = splitmix64(state)
        u.append((z >> 11) * 2.0 ** -53)
    return 0.2 + 0.4 * u[0], 0.3 + 0.5 * u[1], 0.004 + 0.006 * u[2]


def ensemble2d(n_scenes=512, seed=20240601, lx=1.61, lz=1.0, gamma=7.0, diffusion=DIFF_ANTUONO, layers=3, **kw):
    """Batch of independent 2-D dam breaks; scene k uses ensemble_draw(seed + k)."""
    scenes = []
    for k in range(n_scenes):
        w, H, dx = ensemble_draw(seed + k)
        scenes.append(Case.dambreak(2, dx, lx, 0.0, lz, w, H, layers=layers, gamma=gamma, cfl=0.2,
                                    diffusion=diffusion, **kw))
    return Batch(scenes)
