"""Pins of the oracle's lattice constants against what the mathematics fixes.

The tables the oracle transcribes from the reference (direction set SP:183-187,
weights SP:195-197, LR SP:85, M SP:64-82, inv_M SP:83/110) are checked here by
properties that a transcription slip would break, not by retyping them:
  * e is the D3Q19 set with opposite directions adjacent and LR[s] = opp(s);
  * w gives the isotropic moments of a D3Q19 lattice with c_s^2 = 1/3;
  * every row of M equals a polynomial in e (the MRT moment basis);
  * inv_M is the exact rational inverse of M (computed here with Fractions),
    rounded once, and M w = e_0.
"""
from fractions import Fraction
from itertools import product

import numpy as np
import pytest

from oracle import COracle
from oracle import oracle_np as onp

# The moment basis as polynomials of the direction vector (x, y, z).
def _n2(x, y, z):
    return x * x + y * y + z * z


BASIS = [
    lambda x, y, z: 1,
    lambda x, y, z: _n2(x, y, z) - 1,
    lambda x, y, z: 3 * _n2(x, y, z) ** 2 - 6 * _n2(x, y, z) + 1,
    lambda x, y, z: x,
    lambda x, y, z: (3 * _n2(x, y, z) - 5) * x,
    lambda x, y, z: y,
    lambda x, y, z: (3 * _n2(x, y, z) - 5) * y,
    lambda x, y, z: z,
    lambda x, y, z: (3 * _n2(x, y, z) - 5) * z,
    lambda x, y, z: 3 * x * x - _n2(x, y, z),
    lambda x, y, z: (2 * _n2(x, y, z) - 3) * (3 * x * x - _n2(x, y, z)),
    lambda x, y, z: y * y - z * z,
    lambda x, y, z: (2 * _n2(x, y, z) - 3) * (y * y - z * z),
    lambda x, y, z: x * y,
    lambda x, y, z: y * z,
    lambda x, y, z: x * z,
    lambda x, y, z: (y * y - z * z) * x,
    lambda x, y, z: (z * z - x * x) * y,
    lambda x, y, z: (x * x - y * y) * z,
]


def test_direction_set():
    e = onp.E
    want = {(0, 0, 0)} | {v for v in product((-1, 0, 1), repeat=3) if sum(abs(c) for c in v) in (1, 2)}
    assert len(want) == 19
    assert {tuple(int(c) for c in r) for r in e} == want
    assert tuple(e[0]) == (0, 0, 0)
    for s in range(1, 19):
        assert np.all(e[onp.LR[s]] == -e[s])
        # opposite directions are adjacent: (1,2), (3,4), ...
        assert onp.LR[s] == (s + 1 if s % 2 else s - 1)


def test_weights_isotropy():
    e, w = onp.E.astype(float), onp.W
    assert abs(w.sum() - 1) < 1e-15
    assert np.allclose(np.einsum("q,qa->a", w, e), 0, atol=1e-16)
    assert np.allclose(np.einsum("q,qa,qb->ab", w, e, e), np.eye(3) / 3, atol=1e-15)
    assert np.allclose(np.einsum("q,qa,qb,qc->abc", w, e, e, e), 0, atol=1e-16)
    d = np.eye(3)
    iso4 = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d)) / 9
    assert np.allclose(np.einsum("q,qa,qb,qc,qd->abcd", w, e, e, e, e), iso4, atol=1e-15)


@pytest.mark.parametrize("src", ["numpy", "c64", "c32"])
def test_M_is_polynomial_basis(src):
    want = np.array([[b(*map(int, v)) for v in onp.E] for b in BASIS], float)
    if src == "numpy":
        got = onp.M
    else:
        o = COracle(2, 2, 2, real="f64" if src == "c64" else "f32")
        o.init_simulation()
        got = o.M()
    np.testing.assert_array_equal(got, want)


def _exact_inverse(M):
    n = M.shape[0]
    a = [[Fraction(int(M[r, c])) for c in range(n)] + [Fraction(int(r == c)) for c in range(n)] for r in range(n)]
    for c in range(n):
        p = next(r for r in range(c, n) if a[r][c] != 0)
        a[c], a[p] = a[p], a[c]
        piv = a[c][c]
        a[c] = [x / piv for x in a[c]]
        for r in range(n):
            if r != c and a[r][c] != 0:
                m = a[r][c]
                a[r] = [x - m * y for x, y in zip(a[r], a[c])]
    return [[a[r][n + c] for c in range(n)] for r in range(n)]


def test_inverse_M_exact():
    inv = _exact_inverse(onp.M)
    nz = [(r, c) for r in range(19) for c in range(19) if inv[r][c] != 0]
    assert len(nz) == 207  # as many nonzeros as M itself
    dens = {inv[r][c].denominator for r, c in nz}
    assert dens <= {1, 2, 3, 4, 6, 8, 12, 16, 18, 24, 36, 48, 72}
    for real, tol in (("f64", 1e-15), ("f32", 0.0)):
        o = COracle(2, 2, 2, real=real)
        o.init_simulation()
        got = o.invM()
        for r, c in nz:
            exact = float(inv[r][c])
            if real == "f32":
                # f32(LAPACK-like double inverse) == f32(exact rational) on every true nonzero
                assert got[r, c] == float(np.float32(exact))
            else:
                assert abs(got[r, c] - exact) <= tol * max(1.0, abs(exact))
        zeros = np.array([[inv[r][c] == 0 for c in range(19)] for r in range(19)])
        assert np.abs(got[zeros]).max() < 1e-15  # only elimination noise where the exact entry is 0
        np.testing.assert_allclose(got @ onp.M, np.eye(19), atol=1e-6 if real == "f32" else 1e-14)


def test_M_w_is_e0_and_gram_blocks():
    np.testing.assert_array_equal(onp.M @ (onp.W * 36), np.r_[36, np.zeros(18)])
    G = onp.M @ onp.M.T
    blocks = [(0, 1, 2), (3, 4), (5, 6), (7, 8), (9, 10), (11, 12)]
    mask = np.eye(19, dtype=bool)
    for b in blocks:
        for i in b:
            for j in b:
                mask[i, j] = True
    assert np.all(G[~mask] == 0)


def test_relaxation_vector_and_magic_parameter():
    for niu in (0.01, 0.1, 0.16667, 1.0, 3.0):
        for conv in ("reference", "documented"):
            S = onp.relaxation(niu, conv)
            assert S[0] == S[3] == S[5] == S[7] == 0
            s_v, s_o = S[1], S[4]
            tau = niu / 3 + 0.5 if conv == "reference" else 3 * niu + 0.5
            assert abs(s_v - 1 / tau) < 1e-15
            # (1/s_v - 1/2)(1/s_o - 1/2) = 3/16 for every viscosity (SP:129)
            assert abs((1 / s_v - 0.5) * (1 / s_o - 0.5) - 3 / 16) < 1e-12
            assert list(np.nonzero(S == s_o)[0]) == [4, 6, 8, 16, 17, 18] or s_o == s_v


def test_feq_moments():
    rng = np.random.default_rng(0)
    for _ in range(10):
        rho = 1 + 0.1 * rng.standard_normal()
        u = 0.1 * rng.standard_normal(3)
        f = onp.feq(np.array(rho), u)
        e = onp.E.astype(float)
        assert abs(f.sum() - rho) < 1e-14
        np.testing.assert_allclose(f @ e, rho * u, atol=1e-15)
        np.testing.assert_allclose(np.einsum("q,qa,qb->ab", f, e, e), rho * (np.outer(u, u) + np.eye(3) / 3),
                                   atol=1e-14)


def test_meq_is_moments_of_incompressible_feq():
    """meq_vec(rho, u) = M feq(1, u) in every moment but m0, which is rho (SP:209-215)."""
    rng = np.random.default_rng(1)
    for _ in range(10):
        rho = 1 + 0.1 * rng.standard_normal()
        u = 0.1 * rng.standard_normal(3)
        meq = onp.meq_vec(rho, u[None])[0]
        mf = onp.M @ onp.feq(np.array(1.0), u)
        assert meq[0] == rho
        np.testing.assert_allclose(meq[1:], mf[1:], atol=1e-15)
