"""Generator of the coefficients of dev::exp2_fp64 (qem_device.h), the FP64
2^f of the FP64 kernels: the degree-10 polynomial interpolating 2^f
at the 11 Chebyshev points of [-1/2, 1/2] (near-minimax, Mason & Handscomb,
"Chebyshev Polynomials", ch. 6), computed in 50-digit arithmetic and
rounded to double, c[0] + c[1] f + ... + c[10] f^10.  Run to print them and
the maximum relative error of the double Horner evaluation on a dense grid
(4.7e-16); tests/test_library_cpu.py checks that qem_device.h holds these
values.
"""
import mpmath as mp

mp.mp.dps = 50
DEG = 10


def coefficients(deg=DEG):
    n = deg + 1
    xs = [mp.mpf(1) / 2 * mp.cos(mp.pi * (2 * i + 1) / (2 * n)) for i in range(n)]
    ys = [mp.power(2, x) for x in xs]
    # Newton divided differences, then expand to monomials
    dd = list(ys)
    for j in range(1, n):
        for i in range(n - 1, j - 1, -1):
            dd[i] = (dd[i] - dd[i - 1]) / (xs[i] - xs[i - j])
    poly = [mp.mpf(0)] * n
    for j in range(n - 1, -1, -1):
        # poly = poly * (x - xs[j]) + dd[j]
        new = [mp.mpf(0)] * n
        for k in range(n - 1):
            new[k + 1] += poly[k]
            new[k] -= poly[k] * xs[j]
        new[0] += dd[j]
        poly = new
    return [float(c) for c in poly]


if __name__ == "__main__":
    import numpy as np
    c = coefficients()
    for i, v in enumerate(c):
        print(f"c{i} = {v!r}")
    f = np.linspace(-0.5, 0.5, 200001)
    p = np.zeros_like(f)
    for v in reversed(c):
        p = p * f + v
    ref = np.array([float(mp.power(2, mp.mpf(x))) for x in f[::200]])
    err = np.max(np.abs(p[::200] - ref) / ref)
    print("max rel err (double Horner):", err)
