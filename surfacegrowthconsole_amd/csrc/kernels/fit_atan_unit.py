"""Fit of atan_unit (sgk_device.h): atan(t) ~ t P(t^2) on [0, 1], P of degree 8 with P(0) = 1.

The mixed form's angle deviation takes atan of a ratio |u/v| <= 1, so only [0, 1] is needed.  With
s = t^2 and g(s) = atan(sqrt(s)) / sqrt(s), the relative error of t P(s) is P(s) / g(s) - 1; P(s) =
1 + s (c_1 + c_2 s + ... + c_8 s^7) keeps atan'(0) = 1 exactly.  The coefficients minimise the
maximum relative error by Lawson's iteration (weighted least squares whose weights grow where the
error is largest; it converges to the minimax fit) on 4,001 Chebyshev points of s in [0, 1], and
are then rounded to 8 significant digits, as written in the header (each then rounds to its float).

    python fit_atan_unit.py        # prints the coefficients, highest degree first, and the error

The error of the FP32 evaluation (Horner with fused multiply-adds, as the device compiles it) over
every float in [2^-14, 1] is checked by tests/test_atan_unit.py.
"""
import numpy as np

DEGREE = 8


def fit(npts=4001, iters=400):
    s = 0.5 - 0.5 * np.cos(np.pi * (np.arange(npts) + 0.5) / npts)  # Chebyshev points of (0, 1)
    g = np.arctan(np.sqrt(s)) / np.sqrt(s)
    # relative error of 1 + s q(s) against g: (1 + s q(s)) / g - 1, linear in the coefficients of q
    A = np.stack([s ** (k + 1) / g for k in range(DEGREE)], axis=1)
    b = 1.0 - 1.0 / g
    w = np.full(npts, 1.0 / npts)
    for _ in range(iters):
        sw = np.sqrt(w)
        c, *_ = np.linalg.lstsq(A * sw[:, None], b * sw, rcond=None)
        err = np.abs(A @ c - b)
        w = w * err
        w /= w.sum()
    return c  # c[k] multiplies s^(k+1)


def coefficients():
    """Highest degree first, ending with the constant 1: the order of the header's Horner scheme."""
    c = fit()
    return [float(f"{x:.8g}") for x in c[::-1]] + [1.0]


if __name__ == "__main__":
    cs = coefficients()
    for x in cs:
        print(f"{x:.8g}")
    t = np.linspace(1e-4, 1.0, 200001)
    p = np.polyval(cs, t * t)
    print("max relative error (FP64 evaluation):", np.max(np.abs(t * p - np.arctan(t)) / np.arctan(t)))
