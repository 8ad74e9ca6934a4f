"""Where the coefficients of mc_log (bb_mc.hip; the same in oracle/mc.c log_f) come from.

mc_log(x) splits x = m 2^e with m in [sqrt(1/2), sqrt(2)) and returns e log 2 + y P(y), y = m - 1, where P is the
degree-9 polynomial that fits log(1 + y) / y in the least-squares sense on 2000 Chebyshev nodes of
[sqrt(1/2) - 1, sqrt(2) - 1] (all ten coefficients free, solved in float64), its coefficients rounded to float32.
The constant term rounds to 1.0 exactly. Running this file prints the coefficients (highest degree first, the order
of the Horner steps in the kernel) and the largest relative error of the float32 evaluation against float64 log over
a sweep of float32 inputs.

    python bbarolo2_amd/csrc/mc_log_fit.py
"""
import numpy as np

NODES = 2000
DEGREE = 9


def coefficients():
    """float32 coefficients c[0..DEGREE] of P(y) = sum c[k] y^k."""
    a, b = np.sqrt(0.5) - 1.0, np.sqrt(2.0) - 1.0
    y = 0.5 * (a + b) + 0.5 * (b - a) * np.cos(np.pi * (np.arange(NODES) + 0.5) / NODES)
    f = np.log1p(y) / y
    A = np.vstack([y ** k for k in range(DEGREE + 1)]).T
    c = np.linalg.lstsq(A, f, rcond=None)[0]
    return c.astype(np.float32)


def evaluate(x, c):
    """The kernel's evaluation in float32 (Horner steps with fused multiply-adds emulated in float64)."""
    x = np.asarray(x, np.float32)
    i = x.view(np.uint32)
    e = ((i >> 23) & 0xFF).astype(np.int64) - 127
    m = ((i & 0x007FFFFF) | 0x3F800000).view(np.float32)
    big = m > np.float32(1.41421356)
    m = np.where(big, m * np.float32(0.5), m).astype(np.float32)
    e = np.where(big, e + 1, e)
    y = (m - np.float32(1.0)).astype(np.float32)
    p = np.full_like(y, c[DEGREE])
    for k in range(DEGREE - 1, -1, -1):
        p = (p.astype(np.float64) * y + np.float64(c[k])).astype(np.float32)
    yp = (y * p).astype(np.float32)
    return (e.astype(np.float64) * np.float64(np.float32(0.693147182)) + yp).astype(np.float32)


if __name__ == "__main__":
    c = coefficients()
    for k in range(DEGREE, -1, -1):
        print(f"c[{k}] = {c[k]:.9e}")
    x = np.exp(np.linspace(-20., 20., 2000001)).astype(np.float32)
    err = np.abs(evaluate(x, c).astype(np.float64) - np.log(x.astype(np.float64)))
    print("max abs error over exp(-20 .. 20):", err.max())
