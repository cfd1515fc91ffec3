"""Independent checks for the oracle, written from the physics of the
dielectric slab rather than from the reference's formulas.

  images_of(z, ...)        image charges of a charge between two dielectric
                           interfaces by repeated reflection (the image-charge
                           method, as src/python_lib/ICM.py builds its layers)
  icm_force(...)           brute-force Coulomb force from those images
  smeared_force(...)       force from a unit charge spread in-plane as a 2-D
                           gaussian (alpha/pi) exp(-alpha s^2): the real-space
                           form of the exp(-k^2/(4 alpha)) filter
  k_energy(...)            per-k energy of the slab from the image-series
                           Green's function (no structure factors, no sorting)
  smeared_potential(...)   potential of the same in-plane gaussian charge
  short_pair_potential_brute / self_potential_brute
                           short-range energy terms: every image's point
                           potential minus its smeared potential
"""
from __future__ import annotations

import numpy as np
from scipy import integrate


def images_of(z, g1, g2, Lz, layers):
    """[(weight, z_image)] for a unit charge at z: reflection in z=0 scales by
    g1 (z -> -z), in z=Lz by g2 (z -> 2Lz - z), applied alternately."""
    lay = {0: (1.0, float(z))}
    for n in range(layers):
        w, zz = lay[-n]
        lay[n + 1] = (g2 * w, 2.0 * Lz - zz)
        w, zz = lay[n]
        lay[-n - 1] = (g1 * w, -zz)
    return list(lay.values())


def icm_force(ri, rj, qi, qj, g1, g2, Lz, layers=100, eps=1.0):
    """Force on charge i from charge j and all of j's images (no xy
    periodicity).  With rj == ri the real charge is skipped (self images)."""
    F = np.zeros(3)
    for w, zimg in images_of(rj[2], g1, g2, Lz, layers):
        d = np.array([ri[0] - rj[0], ri[1] - rj[1], ri[2] - zimg])
        r2 = float(d @ d)
        if r2 == 0.0:
            continue
        F += qi * qj * w / (4.0 * np.pi * eps * r2) * d / np.sqrt(r2)
    return F


def green_series_k(k, zi, zj, g1, g2, Lz, layers=200):
    """1/2 sum over images of w exp(-k |zi - z_img|): the 2-D Fourier kernel
    (times k/(2 pi)) of the potential of j and its images at i."""
    return 0.5 * sum(w * np.exp(-k * abs(zi - zz)) for w, zz in images_of(zj, g1, g2, Lz, layers))


def smeared_force(r, alpha):
    """Force on a unit point charge at r = (x, y, z) from a unit charge
    spread over the plane z = 0 with density (alpha/pi) exp(-alpha s^2)."""
    rx, ry, rz = r
    out = []
    for comp in range(3):
        def f(t, u, comp=comp):
            sx, sy = u * np.cos(t), u * np.sin(t)
            d = np.array([rx - sx, ry - sy, rz])
            return d[comp] / (d @ d) ** 1.5 * (alpha / np.pi) * np.exp(-alpha * u * u) * u
        v, _ = integrate.dblquad(f, 0.0, 8.0 / np.sqrt(alpha), 0.0, 2.0 * np.pi,
                                 epsabs=1e-12, epsrel=1e-10)
        out.append(v)
    return np.array(out)


def short_pair_brute(ri, rj, alpha, g1, g2, Lz, layers=12):
    """Short-range force on i from j: every image of j contributes its point
    Coulomb force minus its in-plane gaussian-smeared force (1/(4 pi))."""
    F = np.zeros(3)
    for w, zimg in images_of(rj[2], g1, g2, Lz, layers):
        if abs(w) < 1e-15:
            continue
        r = np.array([ri[0] - rj[0], ri[1] - rj[1], ri[2] - zimg])
        F += w * (r / np.linalg.norm(r) ** 3 - smeared_force(r, alpha)) / (4.0 * np.pi)
    return F


def smeared_potential(r, alpha):
    """Potential at r = (x, y, z) of a unit charge spread over the plane
    z = 0 with density (alpha/pi) exp(-alpha s^2) (no 1/(4 pi) factor)."""
    rx, ry, rz = r

    def f(t, u):
        sx, sy = u * np.cos(t), u * np.sin(t)
        d2 = (rx - sx) ** 2 + (ry - sy) ** 2 + rz * rz
        return (alpha / np.pi) * np.exp(-alpha * u * u) * u / np.sqrt(d2)
    v, _ = integrate.dblquad(f, 0.0, 8.0 / np.sqrt(alpha), 0.0, 2.0 * np.pi,
                             epsabs=1e-13, epsrel=1e-11)
    return v


def short_pair_potential_brute(ri, rj, alpha, g1, g2, Lz, layers=12):
    """Short-range pair energy of unit charges i, j: every image of j (the
    charge itself included) contributes point minus smeared potential at i,
    over 4 pi (eps_0 = 1)."""
    u = 0.0
    for w, zimg in images_of(rj[2], g1, g2, Lz, layers):
        if abs(w) < 1e-15:
            continue
        r = np.array([ri[0] - rj[0], ri[1] - rj[1], ri[2] - zimg])
        u += w * (1.0 / np.linalg.norm(r) - smeared_potential(r, alpha)) / (4.0 * np.pi)
    return u


def self_potential_brute(ri, alpha, g1, g2, Lz, layers=12):
    """Short-range self energy of a unit charge with its own images (the
    charge itself excluded): 1/2 sum of point minus smeared potentials."""
    s = 0.0
    lay = images_of(ri[2], g1, g2, Lz, layers)
    for w, zimg in lay[1:]:          # lay[0] is the charge itself
        if abs(w) < 1e-15:
            continue
        r = np.array([0.0, 0.0, ri[2] - zimg])
        s += w * (1.0 / abs(r[2]) - smeared_potential(r, alpha)) / (4.0 * np.pi)
    return 0.5 * s


def k_energy(X, q, k3, g1, g2, Lz):
    """Energy of one k mode (times Lx Ly eps_0) from the image-series Green's
    function G = 1/2 [e^{-k|dz|} + R / (1 - g1 g2 e^{-2kLz})] with R the
    reflected terms; pair term cos(k . (rho_i - rho_j)) G / k, self term
    1/2 q^2 R_ii / (2k (1 - g1 g2 e^{-2kLz}))."""
    kx, ky, k = k3
    th = kx * X[:, 0] + ky * X[:, 1]
    z = X[:, 2]
    x = g1 * g2 * np.exp(-2.0 * k * Lz)
    n = len(q)
    E = 0.0
    for i in range(n):
        Rs = (g1 * np.exp(-2.0 * k * z[i]) + g2 * np.exp(-2.0 * k * (Lz - z[i]))
              + 2.0 * g1 * g2 * np.exp(-2.0 * k * Lz))
        E += 0.5 * q[i] * q[i] * Rs / (1.0 - x) / (2.0 * k)
        for j in range(i + 1, n):
            dz = z[i] - z[j]
            zp = z[i] + z[j]
            R = (g1 * np.exp(-k * zp) + g2 * np.exp(-k * (2.0 * Lz - zp))
                 + g1 * g2 * (np.exp(-k * (2.0 * Lz + dz)) + np.exp(-k * (2.0 * Lz - dz))))
            E += q[i] * q[j] * np.cos(th[i] - th[j]) * (np.exp(-k * abs(dz)) + R / (1.0 - x)) / (2.0 * k)
    return E


def fd_gradient(fun, X, h=1e-6):
    G = np.zeros_like(X)
    for i in range(X.shape[0]):
        for d in range(X.shape[1]):
            Xp = X.copy(); Xp[i, d] += h
            Xm = X.copy(); Xm[i, d] -= h
            G[i, d] = (fun(Xp) - fun(Xm)) / (2.0 * h)
    return G


def neighbor_pairs_numpy(xyz, Lx, Ly, rc):
    """Vectorised O(N^2) minimum-image pair set, i < j."""
    x, y = xyz[:, 0], xyz[:, 1]
    dx = x[:, None] - x[None, :]
    dy = y[:, None] - y[None, :]
    dx = dx - Lx * np.rint(dx / Lx)
    dy = dy - Ly * np.rint(dy / Ly)
    d2 = dx * dx + dy * dy
    i, j = np.nonzero(np.triu(d2 <= rc * rc, k=1))
    return set(zip(i.tolist(), j.tolist()))
