"""Shared test helpers: run the oracle and the GPU path on the same cloth."""
import os

import numpy as np

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def square8k_mesh():
    d = np.load(os.path.join(GOLDEN, "square8k_mesh.npz"))
    return d["x"], d["faces"]


def oracle_sim(cloth, params, fp64=False, energy_bending=False, threads=1):
    m = O.Mesh(cloth.x, cloth.faces)
    s = O.Sim(m, cloth.mass, cloth.fixed, implicit_dt=params.implicit_dt, num_substep=params.num_substep,
              iter_count=params.constraint_iter_count, k_stretch=params.stiffness_stretch_spring,
              k_bending=params.stiffness_quadratic_bending, damping=params.damping_cloth,
              gravity=params.gravity, fp64=fp64, energy_bending=energy_bending, threads=threads)
    s._mesh_ref = m
    return s


def rel_maxnorm(a, b):
    """max |a - b| / max |b| over all coordinates."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))
