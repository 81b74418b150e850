"""Generate exact-rational WENO fixtures (tests/golden/weno_fixtures.json).

The reference's weno3/weno5 (core/weno.f90:1-54) are evaluated here in
exact rational arithmetic (fractions.Fraction) with the constants the
reference actually uses: REAL(4) literals widened to double (the f2py build
has no -fdefault-real-8) and tau5 rounded to binary32 (it is an implicitly
typed REAL(4), core/weno.f90:45).  The exact result is rounded to the
nearest double once.  A floating-point evaluation of the same formula is
then expected within a few ulps of the input scale.

Inputs are seeded; cases cover random, smooth, linear, constant, step and
zero stencils.  Run:  python tests/golden/gen_weno_fixtures.py
"""
import json
import os
import struct
from fractions import Fraction as Fr

import numpy as np


def f32(x):
    """Nearest binary32 value of a rational or float, as a Fraction."""
    return Fr(struct.unpack("f", struct.pack("f", float(x)))[0])


def f32_of_fraction(q):
    # float(q) rounds the exact rational to the nearest double; the double
    # is then rounded to binary32 (double rounding is immaterial here)
    return f32(float(q))


def lit(num, den=1):
    """A REAL(4) literal num./den. folded in binary32 then widened."""
    return f32(np.float32(num) / np.float32(den))


C13, C76, C116, C16, C56 = lit(1, 3), lit(7, 6), lit(11, 6), lit(1, 6), lit(5, 6)
K1, K2 = lit(13, 12), Fr(1, 4)
EPS5 = f32(np.float32(1e-16))
EPS3 = f32(np.float32(1e-14))


def weno5_exact(qmm, qm, q0, qp, qpp):
    qmm, qm, q0, qp, qpp = map(Fr, (qmm, qm, q0, qp, qpp))
    qi1 = C13 * qmm - C76 * qm + C116 * q0
    qi2 = -C16 * qm + C56 * q0 + C13 * qp
    qi3 = C13 * q0 + C56 * qp - C16 * qpp
    b1 = K1 * (qmm - 2 * qm + q0) ** 2 + K2 * (qmm - 4 * qm + 3 * q0) ** 2
    b2 = K1 * (qm - 2 * q0 + qp) ** 2 + K2 * (qm - qp) ** 2
    b3 = K1 * (q0 - 2 * qp + qpp) ** 2 + K2 * (3 * q0 - 4 * qp + qpp) ** 2
    tau5 = f32_of_fraction(abs(b1 - b3))
    w1 = 1 + tau5 / (b1 + EPS5)
    w2 = 6 * (1 + tau5 / (b2 + EPS5))
    w3 = 3 * (1 + tau5 / (b3 + EPS5))
    return (w1 * qi1 + w2 * qi2 + w3 * qi3) / (w1 + w2 + w3)


def weno3_exact(qm, q0, qp):
    qm, q0, qp = map(Fr, (qm, q0, qp))
    qi1 = (-qm + 3 * q0) / 2
    qi2 = (q0 + qp) / 2
    b1 = (q0 - qm) ** 2
    b2 = (qp - q0) ** 2
    tau = abs(b2 - b1)
    w1 = 1 + tau / (b1 + EPS3)
    w2 = (1 + tau / (b2 + EPS3)) * 2
    return (w1 * qi1 + w2 * qi2) / (w1 + w2)


def cases(rng):
    out5, out3 = [], []
    for _ in range(40):                              # random O(1)
        out5.append(list(rng.standard_normal(5)))
        out3.append(list(rng.standard_normal(3)))
    for s in (1e-3, 1e3):                            # other scales
        for _ in range(10):
            out5.append(list(s * rng.standard_normal(5)))
            out3.append(list(s * rng.standard_normal(3)))
    x = np.linspace(0, 1, 5)
    for k in (0.3, 1.0, 2.5):                        # smooth
        out5.append(list(np.sin(k * x + 0.2)))
        out3.append(list(np.sin(k * x[:3] + 0.2)))
    out5 += [[1.0, 2.0, 3.0, 4.0, 5.0], [-2.0, -1.5, -1.0, -0.5, 0.0]]     # linear
    out5 += [[0.7] * 5, [0.0] * 5]                   # constant, zero
    out5 += [[0.0, 0.0, 0.0, 1.0, 1.0], [1.0, 1.0, 0.0, 0.0, 0.0], [0.0, 0.0, 1.0, 1.0, 1.0]]
    out3 += [[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.5, 0.5, 0.5], [0.0, 0.0, 0.0]]
    return out5, out3


def main():
    rng = np.random.default_rng(20240601)
    in5, in3 = cases(rng)
    data = {
        "description": "exact-rational weno5/weno3 of core/weno.f90 with REAL(4) literals "
                       "and REAL(4) tau5; outputs rounded once to double",
        "weno5": [{"q": q, "out": float(weno5_exact(*q))} for q in in5],
        "weno3": [{"q": q, "out": float(weno3_exact(*q))} for q in in3],
        "constants": {"1/3": float(C13), "7/6": float(C76), "11/6": float(C116),
                      "1/6": float(C16), "5/6": float(C56), "13/12": float(K1),
                      "eps5": float(EPS5), "eps3": float(EPS3)},
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weno_fixtures.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", path, len(data["weno5"]), len(data["weno3"]))


if __name__ == "__main__":
    main()
