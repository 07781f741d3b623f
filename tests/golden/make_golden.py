"""Writes derivative_kats.json: exact known answers for the 1-D derivative
kernels, computed in rational arithmetic (fractions.Fraction).

The reference ships no tests, fixtures or golden vectors, so the fixtures
here come from mathematics the algorithms must satisfy:
  - WENO5-JS: every 3-point candidate stencil differentiates a quadratic
    exactly, so both one-sided derivatives equal p'(x_i) for any weights;
  - ENO3: every 4-point stencil differentiates a cubic exactly.
Each case stores the 7 samples around x_i (spacing h) and p'(x_i).

    python tests/golden/make_golden.py   # rewrites the JSON next to it
"""
import json
import os
import random
from fractions import Fraction as F


def case(coeffs, h, x0):
    xs = [x0 + (k - 3) * h for k in range(7)]
    p = lambda x: sum(c * x ** n for n, c in enumerate(coeffs))
    dp = sum(n * c * x0 ** (n - 1) for n, c in enumerate(coeffs) if n > 0)
    return {"samples": [float(p(x)) for x in xs], "h": float(h), "derivative": float(dp),
            "coeffs": [str(c) for c in coeffs], "x0": str(x0)}


def main():
    rnd = random.Random(20261203)
    out = {"weno5_quadratic": [], "eno3_cubic": []}
    for _ in range(12):
        h = F(1, rnd.choice([16, 32, 64, 128, 256]))
        x0 = F(rnd.randint(-40, 40), 64)
        quad = [F(rnd.randint(-9, 9), rnd.randint(1, 7)) for _ in range(3)]
        cub = [F(rnd.randint(-9, 9), rnd.randint(1, 7)) for _ in range(4)]
        out["weno5_quadratic"].append(case(quad, h, x0))
        out["eno3_cubic"].append(case(cub, h, x0))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "derivative_kats.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
