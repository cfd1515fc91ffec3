"""CPU pins for the oracle's energy (qo_energy with qo_pair_potential and
qo_self_potential), which the adjoint kernels are checked against.

  - short range: the pair energy of two unit charges and the self energy of
    one equal the point-minus-smeared image-charge sums of pins.py (every
    image explicit, 2-D quadrature of the smeared potential);
  - long range: the energy of one k mode equals the image-series energy of
    pins.k_energy (times the batch weight / (Lx Ly eps_0)) plus the k = 0
    term -1/(2 Lx Ly eps_0) sum_{i<j} q_i q_j |z_i - z_j| written out here;
  - derivatives: dE/dq (exact: E is quadratic in q), dE/dgamma and -dE/dx
    against central differences of the oracle's own energy, and
    E = 1/2 sum_i q_i phi_i.
A dropped image, a wrong sign or a swapped index in any of these terms moves
one of the checks by far more than its tolerance.
"""
import numpy as np
import pytest

import oracle as O
import pins
from qem4molly_amd.workloads import rel_rms


def _short_energy(p, X, q):
    return O.energy(p, X, q, np.array([[0, 1]], np.int32), short=True, long=False)[0]


@pytest.mark.parametrize("g1,g2,alpha", [(0.0, 0.0, 1.0), (0.5, -0.4, 1.0), (0.7, 0.6, 2.5)])
def test_short_energy_matches_point_minus_smeared_images(g1, g2, alpha):
    # converged limits (SWAP_LIMITS), tiny accuracy, many nodes, as for the
    # force pin of the same geometry
    ri = np.array([5.0, 5.0, 0.9])
    rj = np.array([5.8, 4.4, 0.5])
    p = O.params((10, 10, 2.0), 3.0, alpha, g1, g2, 1.0, 1e-14, 120, O.SEM_CORRECTED)
    X = np.array([ri, rj])
    s0 = _short_energy(p, X, np.array([1.0, 0.0]))
    s1 = _short_energy(p, X, np.array([0.0, 1.0]))
    u = _short_energy(p, X, np.array([1.0, 1.0])) - s0 - s1
    layers = 40 if (g1 or g2) else 0
    assert u == pytest.approx(pins.short_pair_potential_brute(ri, rj, alpha, g1, g2, 2.0, layers),
                              rel=1e-8, abs=1e-13)
    assert s0 == pytest.approx(pins.self_potential_brute(ri, alpha, g1, g2, 2.0, layers),
                               rel=1e-8, abs=1e-13)
    assert s1 == pytest.approx(pins.self_potential_brute(rj, alpha, g1, g2, 2.0, layers),
                               rel=1e-8, abs=1e-13)


@pytest.mark.parametrize("seed", [0, 1])
def test_long_energy_per_k_matches_image_series(seed):
    rng = np.random.default_rng(seed)
    Lx, Ly, Lz = 8.0, 7.0, 3.0
    n = 7
    X = np.column_stack([rng.uniform(0, Lx, n), rng.uniform(0, Ly, n), rng.uniform(0.2, 2.8, n)])
    q = rng.normal(size=n)
    g1, g2 = 0.6, -0.45
    p = O.params((Lx, Ly, Lz), 2.0, 1.0, g1, g2, 1.0, 1e-6, 30)
    kx, ky = 2 * np.pi * 1 / Lx, 2 * np.pi * -2 / Ly
    k3 = np.array([kx, ky, np.hypot(kx, ky)])
    w = 0.37
    E = O.energy(p, X, q, np.zeros((0, 2), np.int32), kvecs=k3[None, :], weight=w,
                 short=False, long=True)[0]
    A = Lx * Ly * 1.0
    E0 = 0.0
    for i in range(n):
        for j in range(i + 1, n):
            E0 -= q[i] * q[j] * abs(X[i, 2] - X[j, 2]) / (2.0 * A)
    assert E == pytest.approx(E0 + w / A * pins.k_energy(X, q, k3, g1, g2, Lz), rel=1e-12)


def test_energy_derivatives_by_differences():
    rng = np.random.default_rng(7)
    n = 10
    Lx, Ly, Lz = 20.0, 20.0, 4.0
    X = np.column_stack([rng.uniform(7, 12, n), rng.uniform(7, 12, n), rng.uniform(0.5, 3.5, n)])
    q = rng.normal(size=n)
    # every pair inside r_cutoff (no cutoff crossings under the differences);
    # corrected semantics: the force is minus the gradient at frozen limits,
    # and with accuracy 1e-12 the limits' own motion is negligible
    p = O.params((Lx, Ly, Lz), 9.0, 0.8, 0.55, -0.35, 1.0, 1e-12, 40, O.SEM_CORRECTED)
    pairs = O.neighbor_pairs(X, Lx, Ly, 9.0)
    assert len(pairs) == n * (n - 1) // 2
    K = np.array([[2 * np.pi / Lx, 0.0, 2 * np.pi / Lx],
                  [2 * np.pi / Lx, 2 * np.pi * 2 / Ly, 2 * np.pi * np.hypot(1 / Lx, 2 / Ly)],
                  [-2 * np.pi * 3 / Lx, 2 * np.pi / Ly, 2 * np.pi * np.hypot(3 / Lx, 1 / Ly)]])
    wt = 0.8

    def energy(Y=X, qq=q, pp=p):
        return O.energy(pp, Y, qq, pairs, kvecs=K, weight=wt)

    E, phi, dg = energy()
    assert E == pytest.approx(0.5 * np.dot(q, phi), rel=1e-13)
    dq = rng.normal(size=n)
    h = 1e-3
    assert (energy(qq=q + h * dq)[0] - energy(qq=q - h * dq)[0]) / (2 * h) == \
        pytest.approx(np.dot(phi, dq), rel=1e-10)
    for w in range(2):
        hp = O.params((Lx, Ly, Lz), 9.0, 0.8, 0.55 + (1e-6 if w == 0 else 0.0),
                      -0.35 + (1e-6 if w == 1 else 0.0), 1.0, 1e-12, 40, O.SEM_CORRECTED)
        hm = O.params((Lx, Ly, Lz), 9.0, 0.8, 0.55 - (1e-6 if w == 0 else 0.0),
                      -0.35 - (1e-6 if w == 1 else 0.0), 1.0, 1e-12, 40, O.SEM_CORRECTED)
        fd = (energy(pp=hp)[0] - energy(pp=hm)[0]) / 2e-6
        assert dg[w] == pytest.approx(fd, rel=1e-7, abs=1e-10 * abs(E))
    F = O.short_forces(p, X, q, pairs) + O.long_forces(p, X, q, K, weight=wt)
    G = pins.fd_gradient(lambda Y: energy(Y=Y)[0], X, h=1e-6)
    assert rel_rms(-G, F) < 1e-7
