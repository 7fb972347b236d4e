"""Derives the FP32 polynomial coefficients of the fast transcendental paths in
mapx_dev.hiph (mx::atan2_f32, mx::sincos_f32) and prints them as hex floats.

Each polynomial is the minimax (Chebyshev-optimal) fit of the function's
reduced form on its interval, minimising the maximum relative error of the
final result, solved as a linear program over 6000 Chebyshev-spaced samples
(equivalent to the Remez exchange at this size); coefficients are then rounded
to FP32 and the FP32 evaluation (Horner with fused multiply-adds) is measured
in ulps against FP64 on 2e6 points of the interval:

  atan(t) = t + t s P(s),  s = t^2,  t in [0, 1]          (8 coefficients)
  sin(r)  = r + r u S(u),  u = r^2,  |r| <= pi/4 + 1e-3   (3 coefficients)
  cos(r)  = 1 + u C(u),    u = r^2,  |r| <= pi/4 + 1e-3   (4 coefficients)

Run: python map_amd/csrc/device/fit_poly.py (numpy, scipy).
"""
import numpy as np
from scipy.optimize import linprog


def minimax(g, w, xs, n):
    """Coefficients c of sum c_j x^j minimising max |w (sum c_j x^j - g)|."""
    V = np.vander(xs, n, increasing=True)
    A = np.vstack([np.hstack([w[:, None] * V, -np.ones((len(xs), 1))]),
                   np.hstack([-w[:, None] * V, -np.ones((len(xs), 1))])])
    b = np.concatenate([w * g, -w * g])
    c = np.zeros(n + 1)
    c[-1] = 1
    r = linprog(c, A_ub=A, b_ub=b, bounds=[(None, None)] * (n + 1), method="highs")
    return [np.float32(v) for v in r.x[:n]]


def fma32(a, b, c):
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def horner32(cs, u):
    p = np.float32(cs[-1]) * np.ones_like(u)
    for cc in cs[-2::-1]:
        p = fma32(p, u, cc)
    return p


def max_ulp(y, ref):
    sp = np.spacing(np.abs(ref.astype(np.float32))).astype(np.float64)
    return float((np.abs(y.astype(np.float64) - ref) / sp).max())


def fit_atan():
    t = np.cos(np.linspace(0, np.pi / 2, 6000))[::-1]
    t = t[t > 1e-4]
    s = t * t
    cs = minimax((np.arctan(t) - t) / (t * s), t * s / np.arctan(t), s, 8)
    tt = np.linspace(0, 1, 2_000_001).astype(np.float32)[1:]
    ss = tt * tt
    return cs, max_ulp(fma32(tt * ss, horner32(cs, ss), tt), np.arctan(tt.astype(np.float64)))


def fit_sin():
    r = np.linspace(1e-4, np.pi / 4 + 1e-3, 6000)
    u = r * r
    cs = minimax((np.sin(r) - r) / (r * u), r * u / np.sin(r), u, 3)
    rr = np.linspace(0, np.pi / 4 + 1e-3, 2_000_001).astype(np.float32)[1:]
    uu = rr * rr
    return cs, max_ulp(fma32(rr * uu, horner32(cs, uu), rr), np.sin(rr.astype(np.float64)))


def fit_cos():
    r = np.linspace(1e-4, np.pi / 4 + 1e-3, 6000)
    u = r * r
    cs = minimax((np.cos(r) - 1) / u, u / np.cos(r), u, 4)
    rr = np.linspace(0, np.pi / 4 + 1e-3, 2_000_001).astype(np.float32)[1:]
    uu = rr * rr
    return cs, max_ulp(fma32(uu, horner32(cs, uu), np.float32(1)), np.cos(rr.astype(np.float64)))


if __name__ == "__main__":
    for name, fit in (("atan P", fit_atan), ("sin S", fit_sin), ("cos C", fit_cos)):
        cs, err = fit()
        print("%s (FP32 evaluation: max %.2f ulp): %s" % (name, err, ", ".join(float(c).hex() + "f" for c in cs)))
