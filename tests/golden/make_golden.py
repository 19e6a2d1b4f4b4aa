"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

Each case is a small seeded system (inputs) with the oracle's outputs under both
distance contracts: FP64 forces (acc, jerk, pot, snap for order 2), nearest
neighbour (address, ds2) and neighbour lists (count + sorted ids, CSR) for the FP64
and the mixed-precision contract.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402
from sapporo2_amd import ics  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
CASES = [  # name, order, n, seed, eps2, mean neighbours
    ("h4_n2", 1, 2, 101, 1e-4, 1.0),
    ("h4_n3", 1, 3, 102, 0.0, 2.0),
    ("h4_n8", 1, 8, 103, 1e-4, 3.0),
    ("h4_n64", 1, 64, 104, 1e-4, 8.0),
    ("h4_n512", 1, 512, 105, 1e-4, 16.0),
    ("h6_n8", 2, 8, 106, 1e-4, 3.0),
    ("h6_n64", 2, 64, 107, 1e-6, 8.0),
    ("h6_n512", 2, 512, 108, 1e-6, 16.0),
]


def csr(cnt, lst):
    off = np.concatenate([[0], np.cumsum(np.minimum(cnt, lst.shape[1]))]).astype(np.int32)
    flat = np.concatenate([lst[i, :min(c, lst.shape[1])] for i, c in enumerate(cnt)]).astype(np.int32)
    return off, flat


def make(name, order, n, seed, eps2, k):
    m, x, v = ics.plummer(n, seed=seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    a = rng.normal(0, 0.5, (n, 3))
    ids = rng.permutation(n).astype(np.int32) + 1
    h2 = ics.neighbour_h2(x, k) if n > 3 else np.full(n, 10.0)
    r = oracle.force(order, x, v, ids, x, v, m, ids, eps2, ai=a if order == 2 else None,
                     aj=a if order == 2 else None, h2=h2)
    mx = oracle.neighbours_mixed(x, ids, x, ids, h2)
    off, flat = csr(r["ngb_cnt"], r["ngb_list"])
    moff, mflat = csr(mx["ngb_cnt"], mx["ngb_list"])
    d = dict(order=order, eps2=eps2, m=m, x=x, v=v, a=a, ids=ids, h2=h2,
             acc=r["acc"], jrk=r["jrk"], pot=r["pot"],
             nn_addr=r["nn_addr"], nn_ds2=r["nn_ds2"], ngb_cnt=r["ngb_cnt"], ngb_off=off, ngb_ids=flat,
             mx_nn_addr=mx["nn_addr"], mx_nn_ds2=mx["nn_ds2"], mx_ngb_cnt=mx["ngb_cnt"],
             mx_ngb_off=moff, mx_ngb_ids=mflat)
    if order == 2:
        d["snp"] = r["snp"]
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **d)


if __name__ == "__main__":
    for c in CASES:
        make(*c)
        print("wrote", c[0])
