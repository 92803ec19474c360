"""The coarse solve of the two-level pressure CG, restated in numpy.

When A has the constant null vector (the pressure correction without a
fixed-pressure boundary), E = Z^T A Z is singular with every row summing to
zero; the device inverts E + tau 1 1^T (tau = mean diagonal / m) instead.
On the zero-sum vectors s = Z^T r that CG produces this must act as the
pseudo-inverse E^+, and the additive preconditioner B^-1 + Z E^-1 Z^T must
stay definite. A wrong tau sign or a dropped rank-one term fails these."""
import numpy as np


def _laplacian(n, rng):
    # negative semi-definite graph Laplacian with zero row sums, the sign
    # convention of the pressure-correction matrix
    W = np.zeros((n, n))
    for i in range(n - 1):
        W[i, i + 1] = W[i + 1, i] = rng.uniform(0.5, 2.0)
    for _ in range(2 * n):
        i, j = rng.integers(0, n, 2)
        if i != j:
            W[i, j] = W[j, i] = rng.uniform(0.5, 2.0)
    return W - np.diag(W.sum(1))


def test_regularised_inverse_is_pseudo_inverse_on_zero_sum_vectors():
    rng = np.random.default_rng(5)
    n, m = 60, 7
    A = _laplacian(n, rng)
    agg = np.arange(n) % m
    Z = np.zeros((n, m))
    Z[np.arange(n), agg] = 1.0
    E = Z.T @ A @ Z
    assert np.allclose(E.sum(1), 0.0, atol=1e-12)  # singular: rows sum to zero
    tau = np.trace(E) / m / m
    Er = E + tau * np.ones((m, m))
    assert np.all(np.linalg.eigvalsh(Er) < 0)  # definite with the sign of A
    r = rng.standard_normal(n)
    r -= r.mean()  # consistent residual of the singular system
    s = Z.T @ r
    c = np.linalg.solve(Er, s)
    c_pinv = np.linalg.pinv(E) @ s
    assert np.allclose(c, c_pinv, atol=1e-10)
    # and a positive tau (the wrong sign) would not be definite
    assert np.max(np.linalg.eigvalsh(E - tau * np.ones((m, m)))) > 0


def test_additive_two_level_preconditioner_is_definite():
    rng = np.random.default_rng(6)
    n, m = 48, 6
    A = _laplacian(n, rng) - 0.05 * np.eye(n)  # a fixed-pressure boundary: definite
    agg = np.repeat(np.arange(m), n // m)
    Z = np.zeros((n, m))
    Z[np.arange(n), agg] = 1.0
    E = Z.T @ A @ Z
    Binv = np.diag(1.0 / np.diag(A))  # a one-level sweep (Jacobi)
    M = Binv + Z @ np.linalg.inv(E) @ Z.T
    assert np.allclose(M, M.T)
    assert np.all(np.linalg.eigvalsh(M) < 0)
    # the coarse term removes the aggregate-constant error exactly:
    # Z E^-1 Z^T A (Z y) = Z y
    y = rng.standard_normal(m)
    assert np.allclose(Z @ np.linalg.solve(E, Z.T @ (A @ (Z @ y))), Z @ y)
