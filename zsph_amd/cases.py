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
