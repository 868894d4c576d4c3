"""Writes tests/golden/pdmatgen_golden.npz from the reference's own generator.

The vectors come from PDMATGEN compiled from the reference sources
(LIN/pdmatgen.f + LIN/pmatgeninc.f) by oracle/Makefile into oracle/_ref/,
called exactly as oocdag.cpp:298-302 calls it (1x1 grid, whole matrix).
Run in a container where the reference tree is present:

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pdmatgen_golden.npz")


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libpdmatgen_ref.so not built")
    rec = {}
    for (n, nb) in [(4, 2), (17, 5), (64, 16)]:
        rec[f"S_D_{n}_{nb}"] = oracle.ref_pdmatgen(n, nb, "S", "D", 100)
    rec["S_N_17_5"] = oracle.ref_pdmatgen(17, 5, "S", "N", 100)
    rec["N_N_11x7_3"] = oracle.ref_pdmatgen(7, 3, "N", "N", 100, m=11)
    rec["T_N_11x7_3"] = oracle.ref_pdmatgen(7, 3, "T", "N", 100, m=11)
    rec["F_D_9_4"] = oracle.ref_pdmatgen(9, 4, "F", "D", 100)
    rec["S_D_9_4_seed7"] = oracle.ref_pdmatgen(9, 4, "S", "D", 7)
    # sampled entries of the 4096 x 4096, NB = 256 input (oocdag 4096 256 ...)
    big = oracle.ref_pdmatgen(4096, 256, "S", "D", 100)
    rng = np.random.default_rng(100)
    idx = rng.integers(0, 4096, size=(2000, 2))
    idx = np.concatenate([idx, np.array([[i, i] for i in range(0, 4096, 97)] +
                                        [[4095, 4095], [0, 0], [4095, 0], [0, 4095]])])
    rec["S_D_4096_256_idx"] = idx.astype(np.int64)
    rec["S_D_4096_256_val"] = big[idx[:, 0], idx[:, 1]]
    rec["S_D_4096_256_colsum"] = big.sum(axis=0)
    np.savez_compressed(OUT, **rec)
    print("wrote", OUT, {k: v.shape for k, v in rec.items()})


if __name__ == "__main__":
    main()
