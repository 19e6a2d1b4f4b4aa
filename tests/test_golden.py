"""Golden fixtures (tests/golden/*.npz, written by make_golden.py): the oracle must
reproduce them bit for bit (CPU), and the HIP path must match them (GPU)."""
from __future__ import annotations

import glob
import os

import numpy as np
import pytest

import oracle
from _util import check_forces

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def _load(p):
    with np.load(p, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _lists(g, pre=""):
    off, ids = g[pre + "ngb_off"], g[pre + "ngb_ids"]
    return [ids[off[i]:off[i + 1]] for i in range(len(off) - 1)]


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p) for p in GOLD])
def test_oracle_reproduces_golden(path):
    g = _load(path)
    order = int(g["order"])
    ai = g["a"] if order == 2 else None
    r = oracle.force(order, g["x"], g["v"], g["ids"], g["x"], g["v"], g["m"], g["ids"], float(g["eps2"]),
                     ai=ai, aj=ai, h2=g["h2"])
    for k in ("acc", "jrk", "pot") + (("snp",) if order == 2 else ()):
        assert np.array_equal(r[k], g[k]), k
    assert np.array_equal(r["nn_addr"], g["nn_addr"]) and np.array_equal(r["nn_ds2"], g["nn_ds2"])
    assert np.array_equal(r["ngb_cnt"], g["ngb_cnt"])
    for i, l in enumerate(_lists(g)):
        assert np.array_equal(r["ngb_list"][i, :len(l)], l)
    mx = oracle.neighbours_mixed(g["x"], g["ids"], g["x"], g["ids"], g["h2"])
    assert np.array_equal(mx["nn_addr"], g["mx_nn_addr"]) and np.array_equal(mx["nn_ds2"], g["mx_nn_ds2"])
    assert np.array_equal(mx["ngb_cnt"], g["mx_ngb_cnt"])


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p) for p in GOLD])
def test_gpu_matches_golden(path, prec):
    import sapporo2_amd as sa
    g = _load(path)
    order = int(g["order"])
    n = len(g["m"])
    print(f"golden {os.path.basename(path)} prec={prec}", flush=True)
    with sa.Engine(order, prec) as eng:
        eng.set_j_particles(np.arange(n), g["ids"], g["m"], g["x"], g["v"],
                            g["a"] if order == 2 else None)
        eng.set_time(0.0)
        eng.start_grav_calc(n, g["ids"], g["x"], g["v"], float(g["eps2"]), g["h2"],
                            ai=g["a"] if order == 2 else None)
        r = eng.get_grav_results(n, n)
        # a median over fewer than 5 particles is one particle's error: 10x
        tol = (1e-6 if prec == 1 else 1e-12) * (10 if n < 5 else 1)
        check_forces(r, g, order, tol, label=os.path.basename(path))
        pre = "mx_" if prec == 1 else ""
        nnb, dsmin, found = oracle.api_nn(g[pre + "nn_addr"], g[pre + "nn_ds2"], g["ids"])
        assert np.array_equal(r["nnb"], nnb)
        assert np.array_equal(r["dsmin"][found], dsmin[found])
        eng.read_ngb_list()
        for i, l in enumerate(_lists(g, pre)):
            got, cnt, ov = eng.get_ngb_list(i)
            assert cnt == g[pre + "ngb_cnt"][i], i
            if cnt < 256:  # overflowed lists have unspecified contents
                assert ov == 0 and np.array_equal(got, l), i
            else:
                assert ov == 1, i
