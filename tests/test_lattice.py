"""D3Q19 tables and equilibrium: closed-form identities (CPU).

Counterparts of the reference's tests/test_lbm_algorithms.py:39-130 (weight
normalisation, velocity-set symmetry, equilibrium moments) and the
config/core.py:50 assertion, checked against the lattice's defining moments.
"""
from __future__ import annotations

import numpy as np

from pourover_amd.lattice import CX, CY, CZ, E, OPP, Q, W, equilibrium


def test_weights_and_isotropy():
    assert Q == 19
    assert abs(W.sum() - 1.0) < 1e-15                               # config/core.py:50
    Ef = E.astype(np.float64)
    assert np.allclose(W @ Ef, 0.0, atol=1e-16)                      # sum w e = 0
    second = np.einsum("q,qa,qb->ab", W, Ef, Ef)
    assert np.allclose(second, np.eye(3) / 3.0, atol=1e-16)          # c_s^2 delta
    third = np.einsum("q,qa,qb,qc->abc", W, Ef, Ef, Ef)
    assert np.allclose(third, 0.0, atol=1e-16)
    fourth = np.einsum("q,qa,qb,qc,qd->abcd", W, Ef, Ef, Ef, Ef)
    d = np.eye(3)
    iso = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d)) / 9.0
    assert np.allclose(fourth, iso, atol=1e-16)                      # c_s^4 (dd + dd + dd)


def test_velocity_set_matches_reference_order():
    # config/core.py:36-38 order
    assert CX.tolist() == [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]
    assert CY.tolist() == [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1]
    assert CZ.tolist() == [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1]
    assert len({tuple(e) for e in E}) == 19
    assert (np.abs(E).sum(axis=1) <= 2).all()


def test_opposites():
    for q in range(Q):
        assert (E[OPP[q]] == -E[q]).all()
        assert OPP[OPP[q]] == q
        assert W[OPP[q]] == W[q]
    # pairs of the reference ordering (1-2, 3-4, 5-6, 7-10, 8-9, 11-14, 12-13, 15-18, 16-17)
    assert OPP.tolist() == [0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15]
    # the hard-coded table of cpu_backend.py:189 is wrong for this ordering
    cpu_backend_table = [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17]
    assert cpu_backend_table != OPP.tolist()


def test_equilibrium_moments():
    rng = np.random.default_rng(1)
    rho = 1.0 + 0.1 * rng.standard_normal(50)
    u = 0.05 * rng.standard_normal((3, 50))
    f = equilibrium(rho, u)
    Ef = E.astype(np.float64)
    assert np.allclose(f.sum(axis=0), rho, rtol=1e-14)
    assert np.allclose(np.einsum("qa,qn->an", Ef, f), rho * u, rtol=1e-13, atol=1e-16)
    P = np.einsum("qa,qb,qn->abn", Ef, Ef, f)
    expect = rho / 3.0 * np.eye(3)[:, :, None] + rho * np.einsum("an,bn->abn", u, u)
    assert np.allclose(P, expect, rtol=1e-13, atol=1e-16)


def test_equilibrium_at_rest_is_weights():
    f = equilibrium(np.ones(4), np.zeros((3, 4)))
    assert np.allclose(f, W[:, None], rtol=0, atol=1e-16)
