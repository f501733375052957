"""Dam-break workloads of the benchmark configurations (geometry in metres).

* ``dambreak2d``: 2-D tank 1.61 x 0.8, column 0.4 wide x 0.3 high, dx 0.005
  (4800 fluid particles on a lattice + 3 dummy layers).
* ``dambreak3d``: 3-D tank 3.22 x 1.0 x 1.0, column 1.2 x 1.0 x 0.55, dx 0.008
  (1,293,750 fluid particles + 3 dummy layers on the floor and four walls).
  dx 0.016 gives the 2x-coarser 3-D case (~160k fluid particles).
Both use the Wendland C2 kernel with h = 1.3 dx, Tait EOS gamma 7 with
rho0 = 1000 and c0 = 10 sqrt(2 g H), delta 0.1, alpha 0.02, CFL 0.2.
"""
from . import Case, DIFF_ANTUONO


def dambreak2d(dx=0.005, gamma=7.0, diffusion=DIFF_ANTUONO, lx=1.61, lz=0.8, cw=0.4, ch=0.3, layers=3, **kw):
    return Case.dambreak(2, dx, lx, 0.0, lz, cw, ch, layers=layers, gamma=gamma, cfl=0.2,
                         diffusion=diffusion, **kw)


def dambreak3d(dx=0.008, gamma=7.0, diffusion=DIFF_ANTUONO, lx=3.22, wy=1.0, lz=1.0, cw=1.2, ch=0.55, layers=3,
               **kw):
    return Case.dambreak(3, dx, lx, wy, lz, cw, ch, layers=layers, gamma=gamma, cfl=0.2, diffusion=diffusion,
                         **kw)
