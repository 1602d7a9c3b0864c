"""Write golden (points, weights) vectors of the electrode-fitted quadrature.

Run in the development container only (the reference tree is not shipped):

    python tests/golden/make_quadrature_golden.py

It loads the reference's dctorch/quadrature.py by file path (the module needs
only numpy and scipy; the package __init__, which imports SimPEG, is not
executed) and records k0_quadrature's outputs for the distance ranges and
point counts the tests and benchmark surveys use.  The result is
tests/golden/quadrature_golden.npz (data only).
"""
import importlib.util
import os

import numpy as np

REF = "/root/reference/dctorch/quadrature.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "quadrature_golden.npz")

CASES = [  # (r_min, r_max, nky)
    (2.0, 12.0, 6),
    (1.0, 10.0, 11),
    (2.0, 34.0, 12),     # 61 electrodes at 2 m, WS a=2 n=8 spans 34 m
    (2.0, 36.0, 12),
    (5.0, 120.0, 12),
    (5.0, 35.0, 11),
    (5.0, 120.0, 31),
    (2.0, 48.0, 12),     # the 25,088-cell line (eval+grad benchmark): 10 points kept
    (5.0, 180.0, 12),    # the 62,016-cell Gauss-Newton line
    (2.0, 16.0, 12),     # the uneven batch of 16 lines
    (3.0, 24.0, 11),     # the 5,000-cell regression line with the electrode fit
    (2.857142857142854, 17.142857142857146, 6),  # the smoke / suite case
]


def main():
    spec = importlib.util.spec_from_file_location("ref_quadrature", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    out = {}
    for i, (lo, hi, n) in enumerate(CASES):
        pts, wts, info = mod.k0_quadrature(lo, hi, n)
        out[f"case{i}_in"] = np.array([lo, hi, n], float)
        out[f"case{i}_points"] = pts
        out[f"case{i}_weights"] = wts
        out[f"case{i}_err"] = np.array([info["max_rel_err"]])
        print(f"[{lo}, {hi}] nky={n}: n_used={info['n_used']} err={info['max_rel_err']:.2e}")
    np.savez(OUT, ncases=np.array([len(CASES)]), **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
