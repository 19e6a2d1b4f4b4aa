"""Shared helpers for the parity tests (inputs, loading, comparisons)."""
from __future__ import annotations

import numpy as np

import oracle
from sapporo2_amd import ics

NGB_CAP = 256


def plummer_system(n, seed=1, mean_ngb=32.0):
    m, x, v = ics.plummer(n, seed=seed)
    ids = np.arange(n, dtype=np.int32)
    h2 = ics.neighbour_h2(x, mean_ngb)
    return m, x, v, ids, h2


def random_derivs(n, seed, scale=1.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    return [rng.normal(0.0, scale, (n, 3)) for _ in range(4)]  # acc, jrk, snp, crk


def load_engine(eng, m, x, v, ids, acc=None, jrk=None, snp=None, crk=None, tj=0.0):
    n = len(m)
    eng.set_j_particles(np.arange(n), ids, m, x, v, acc, jrk, snp, crk, tj=tj)


def check_forces(res, ref, order, tol_med, tol_p90=None, label=""):
    """Median (and 90th percentile) per-particle relative errors, as the reference's
    accuracy test prints them (test_gravity_block.cpp:359-362)."""
    out = {}
    keys = ["acc", "jrk", "pot"] + (["snp"] if order >= 2 else [])
    for k in keys:
        e = oracle.rel_err(res[k], ref[k])
        out[k] = (float(np.median(e)), float(np.percentile(e, 90)), float(e.max()))
    for k in ["acc", "pot"]:
        med, p90, _ = out[k]
        assert med <= tol_med, f"{label} {k} median rel err {med:.3e} > {tol_med:.1e} ({out})"
        if tol_p90 is not None:
            assert p90 <= tol_p90, f"{label} {k} p90 rel err {p90:.3e} > {tol_p90:.1e} ({out})"
    return out


def expected_nn(order_prec_f64, xi, ids_i, xj, ids_j, h2):
    """Oracle NN + lists under the working-precision contract."""
    if order_prec_f64:
        r = oracle.force(1, xi, np.zeros_like(xi), ids_i, xj, np.zeros_like(xj),
                         np.ones(len(xj)), ids_j, 1.0, h2=h2)
        return r["nn_addr"], r["nn_ds2"], r["ngb_cnt"], r["ngb_list"]
    r = oracle.neighbours_mixed(xi, ids_i, xj, ids_j, h2)
    return r["nn_addr"], r["nn_ds2"], r["ngb_cnt"], r["ngb_list"]


def check_nn_lists(eng, res, f64, xi, ids_i, xj, ids_j, h2, check_lists=True, label=""):
    """Bit-exact nearest neighbour and neighbour lists against the oracle."""
    nn_addr, nn_ds2, cnt, lst = expected_nn(f64, xi, ids_i, xj, ids_j, h2)
    nnb, dsmin, found = oracle.api_nn(nn_addr, nn_ds2, ids_j)
    assert np.array_equal(res["nnb"], nnb), f"{label} nnb mismatches: {np.sum(res['nnb'] != nnb)}"
    got = res["dsmin"]
    assert np.array_equal(np.isnan(got), ~found), f"{label} dsmin written-mask mismatch"
    assert np.array_equal(got[found], dsmin[found]), \
        f"{label} dsmin mismatches: {np.sum(got[found] != dsmin[found])}"
    if not check_lists or h2 is None:
        return
    ov = eng.read_ngb_list()
    assert ov == int(np.any(cnt >= NGB_CAP)), f"{label} overflow flag"
    bad = 0
    for i in range(len(ids_i)):
        l, n, o = eng.get_ngb_list(i, NGB_CAP)
        if n != cnt[i]:
            bad += 1
            continue
        if cnt[i] < NGB_CAP and not np.array_equal(l, lst[i, :cnt[i]]):
            bad += 1
    assert bad == 0, f"{label} neighbour-list mismatches: {bad}"
