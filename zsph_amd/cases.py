"""Dam-break workloads of the benchmark configurations (geometry in metres).

* ``dambreak2d``: 2-D tank 1.61 x 0.8, column 0.4 wide x 0.3 high, dx 0.005
  (4800 fluid particles on a lattice + 3 dummy layers).
* ``dambreak3d``: 3-D tank 3.22 x 1.0 x 1.0, column 1.2 x 1.0 x 0.55, dx 0.008
  (1,293,750 fluid particles + 3 dummy layers on the floor and four walls).
  dx 0.016 gives the 2x-coarser 3-D case (~160k fluid particles).
* ``wavetank3d``: piston wavemaker and a floating box: tank 2.0 x 0.4 x 0.6,
  water depth 0.35, box 0.2 x 0.2 x 0.16 at half its density (draft 0.08),
  piston stroke 0.03 at 1 Hz, dx 0.01 (~260k particles).
All use the Wendland C2 kernel with h = 1.3 dx, Tait EOS gamma 7 with
rho0 = 1000 and c0 = 10 sqrt(2 g H), delta 0.1, alpha 0.02, CFL 0.2.
"""
import numpy as np

from . import Case, DIFF_ANTUONO, rigid_state


def dambreak2d(dx=0.005, gamma=7.0, diffusion=DIFF_ANTUONO, lx=1.61, lz=0.8, cw=0.4, ch=0.3, layers=3, **kw):
    return Case.dambreak(2, dx, lx, 0.0, lz, cw, ch, layers=layers, gamma=gamma, cfl=0.2,
                         diffusion=diffusion, **kw)


def dambreak3d(dx=0.008, gamma=7.0, diffusion=DIFF_ANTUONO, lx=3.22, wy=1.0, lz=1.0, cw=1.2, ch=0.55, layers=3,
               **kw):
    return Case.dambreak(3, dx, lx, wy, lz, cw, ch, layers=layers, gamma=gamma, cfl=0.2, diffusion=diffusion,
                         **kw)


def wavetank3d(dx=0.01, gamma=7.0, diffusion=DIFF_ANTUONO, lx=2.0, wy=0.4, lz=0.6, depth=0.35,
               box=((0.8, 0.1, 0.27), (1.0, 0.3, 0.43)), layers=3, piston_travel=0.06, **kw):
    """Wave tank with a piston and a floating box (lattice-aligned box faces)."""
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
