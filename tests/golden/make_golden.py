"""Writes the committed residual fixtures fr2d_*.npz from the independent numpy restatement
(tests/np_reference.py).  Inputs: seeded perturbed isentropic vortex states in the reference
layout; outputs: du of example/shock-vortex.jl:26-118 for each (grid, degree, flux, bc)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "fluxreconstruction.jl_amd")]

import np_reference as NR  # noqa: E402
from frhip import FRPSpace2D, ics  # noqa: E402

CASES = [(3, 2, 1, "hll", "ghost"), (4, 3, 2, "rusanov", "periodic"), (5, 4, 3, "hll", "periodic"),
         (3, 3, 4, "rusanov", "ghost"), (6, 2, 3, "rusanov", "ghost")]

for nx, ny, deg, flux, bc in CASES:
    L = 10.0
    ps = FRPSpace2D(0.0, L, nx, 0.0, L, ny, deg, 1, 1)
    u = ics.perturbed(ics.isentropic_vortex(ps, 1.4), seed=100 + deg, amplitude=0.01)
    dx, dy = L / nx, L / ny
    du = NR.fr2d_rhs(u, NR.operators(deg), 1.4, 1 / (dx / 2), 1 / (dy / 2), flux, bc == "periodic")
    name = f"fr2d_{nx}x{ny}_p{deg}_{flux}_{bc}.npz"
    np.savez_compressed(os.path.join(HERE, name), u=u, du=du, dims=np.array([nx, ny, deg]), dx=dx, dy=dy,
                        flux=np.array(flux), bc=np.array(bc))
    print(name)
