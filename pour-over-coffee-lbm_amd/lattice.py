"""D3Q19 lattice constants (host side).

Order and weights of config/core.py:36-47 (reference); opposite directions
by the search of src/core/legacy/lbm_solver.py:431-439 (e_opp = -e).
"""
from __future__ import annotations

import numpy as np

Q = 19
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0], dtype=np.int32)
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1], dtype=np.int32)
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1], dtype=np.int32)
E = np.stack([CX, CY, CZ], axis=1)
W = np.array([1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12, dtype=np.float64)
CS2 = 1.0 / 3.0


def _opposites() -> np.ndarray:
    opp = np.empty(Q, dtype=np.int32)
    for q in range(Q):
        hit = np.nonzero((E == -E[q]).all(axis=1))[0]
        opp[q] = hit[0]
    return opp


OPP = _opposites()


def equilibrium(rho, u) -> np.ndarray:
    """f_eq[q, ...] = w_q rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u.u) in FP64.

    rho: array [...], u: array [3, ...].  Form of lbm_algorithms.py:183-218.
    """
    rho = np.asarray(rho, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    eu = np.tensordot(E.astype(np.float64), u, axes=([1], [0]))
    usq = (u * u).sum(axis=0)
    return W.reshape((Q,) + (1,) * rho.ndim) * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq)
