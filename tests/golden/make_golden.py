"""Writes the golden fixtures of tests/golden/.

pinv_q1_4.npz: the refit operator M = pinv(B)^T for q = 1..4, where B is the
packed-triangle Chebyshev basis at the stretched nodes (cheb.h:166-196),
computed with numpy.linalg.pinv (a library SVD routine, independent of both
the oracle's normal-equations solve and the engine's Householder QR).
Nodes and T_k come from their closed forms, not from either library.

Run: python tests/golden/make_golden.py
"""
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def stretched_nodes(q):
    d = q + 1
    scal = 1.0 / math.cos(0.5 * math.pi / d)
    return np.array([-math.cos((i + 0.5) * math.pi / d) * scal * 0.5 + 0.5 for i in range(d)])


def basis(q):
    d = q + 1
    x = stretched_nodes(q) * 2.0 - 1.0
    T = np.polynomial.chebyshev.chebvander(x, q)
    idx = np.arange(d ** 3)
    cols = []
    for i in range(d):
        for j in range(d - i):
            for k in range(d - i - j):
                cols.append(T[idx % d, k] * T[(idx // d) % d, j] * T[idx // (d * d), i])
    return np.stack(cols, axis=1)


def main():
    out = {f"q{q}": np.linalg.pinv(basis(q)).T for q in range(1, 5)}
    np.savez(os.path.join(HERE, "pinv_q1_4.npz"), **out)


if __name__ == "__main__":
    main()
