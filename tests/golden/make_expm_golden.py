"""Writes expm_golden.json: exp(coeff * Z) for random traceless anti-hermitian
Z, computed with scipy.linalg.expm (scaling-and-squaring Pade, an independent
library routine).  These vectors pin exp_su3_luscher (src/su3/utils.cpp:145-193)
in the oracle and on the device.  Run from the repository root:
    python tests/golden/make_expm_golden.py
"""
import json
import os

import numpy as np
from scipy.linalg import expm


def random_su3_algebra(rng):
    A = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    X = 0.5 * (A - A.conj().T)
    return X - np.trace(X) / 3.0 * np.eye(3)


def main():
    rng = np.random.default_rng(20260219)
    cases = []
    # Frobenius norms of coeff*Z from 1e-4 to ~1.4 (t = |X|^2/2 <= 1)
    norms = np.concatenate([np.logspace(-4, 0, 56), [1.1, 1.2, 1.3, 1.35, 1.38, 1.4, 1.41, 1.414]])
    for k, nrm in enumerate(norms):
        Z = random_su3_algebra(rng)
        coeff = [1.0, 0.25, 0.02, 0.005][k % 4]
        Z *= nrm / (coeff * np.linalg.norm(Z))
        E = expm(coeff * Z)
        cases.append({
            "Z": np.stack([Z.real, Z.imag], -1).reshape(-1).tolist(),
            "coeff": coeff,
            "expm": np.stack([E.real, E.imag], -1).reshape(-1).tolist(),
        })
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "expm_golden.json")
    with open(out, "w") as f:
        json.dump({"source": "scipy.linalg.expm", "cases": cases}, f, indent=0)
    print(f"wrote {len(cases)} cases to {out}")


if __name__ == "__main__":
    main()
