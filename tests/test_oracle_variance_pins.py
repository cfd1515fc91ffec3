"""CPU pins for the oracle's random-batch variance (qo_long_variance), which
the GPU batch-variance pass is checked against.

  - formula: var_i equals P times the two-pass sample variance (numpy, ddof=1)
    of the per-k forces X_ip, each taken from the oracle's long-range force
    of a one-k batch minus its k = 0 term (a different accumulation path);
  - meaning: over many batches drawn as the reference draws them (k from the
    K set with probability exp(-k^2/(4 alpha)) / sum_K, with replacement,
    src/Force_long.jl:234-242), the batch force averages to the full k sum and
    the batch estimate var_i averages to the measured mean |F_i - F_full,i|^2.
A missing factor P, a P in place of P - 1, a wrong batch weight or a k = 0
term left in moves one of these checks by far more than its tolerance.
"""
import numpy as np
import pytest

import oracle as O


def _system(seed, n, L):
    rng = np.random.default_rng(seed)
    X = np.column_stack([rng.uniform(0, L[0], n), rng.uniform(0, L[1], n),
                         rng.uniform(0.2, L[2] - 0.2, n)])
    q = rng.normal(size=n)
    return X, q - q.mean()


@pytest.mark.parametrize("g1,g2", [(0.0, 0.0), (0.6, -0.45)])
def test_variance_formula_two_pass(g1, g2):
    L = (8.0, 7.0, 3.0)
    X, q = _system(1, 9, L)
    p = O.params(L, 2.0, 1.0, g1, g2, 1.0, 1e-4, 20)
    K, prob, sK, _ = O.kset(L[0], L[1], 1.0, 1e-4)
    rng = np.random.default_rng(2)
    P = 6
    kv = K[rng.choice(len(K), P)]
    w = sK / P
    k0 = O.long_forces(p, X, q, np.zeros((0, 3)), weight=w)
    Xp = np.stack([O.long_forces(p, X, q, kv[t:t + 1], weight=w) - k0 for t in range(P)])
    ref = P * np.var(Xp, axis=0, ddof=1).sum(axis=1)
    got = O.long_variance(p, X, q, kv, w)
    assert np.all(ref > 0)
    np.testing.assert_allclose(got, ref, rtol=1e-10)
    # the batch force is k0 + the sum of the X_ip
    np.testing.assert_allclose(O.long_forces(p, X, q, kv, weight=w), k0 + Xp.sum(axis=0),
                               rtol=1e-12, atol=1e-14)


def test_variance_estimates_spread_of_batch_force():
    L = (8.0, 8.0, 3.0)
    X, q = _system(3, 12, L)
    alpha, acc = 1.0, 1e-4
    p = O.params(L, 2.0, alpha, 0.5, -0.3, 1.0, acc, 20)
    K, prob, sK, _ = O.kset(L[0], L[1], alpha, acc)
    pr = prob.astype(np.float64) / prob.sum()
    full = O.long_forces(p, X, q, K, full_sum=True)
    rng = np.random.default_rng(4)
    P, M = 4, 3000
    F = np.empty((M,) + X.shape)
    V = np.empty((M, X.shape[0]))
    for m in range(M):
        kv = K[rng.choice(len(K), P, p=pr)]
        F[m] = O.long_forces(p, X, q, kv, weight=sK / P)
        V[m] = O.long_variance(p, X, q, kv, sK / P)
    # unbiased: the batch mean converges to the full sum
    se = F.std(axis=0) / np.sqrt(M)
    assert np.all(np.abs(F.mean(axis=0) - full) < 5 * se + 1e-12)
    # calibrated: mean batch estimate = measured spread about the full sum
    spread = ((F - full) ** 2).sum(axis=2).mean(axis=0)
    est = V.mean(axis=0)
    assert est.sum() == pytest.approx(spread.sum(), rel=0.1)
    assert np.median(est / spread) == pytest.approx(1.0, rel=0.15)
