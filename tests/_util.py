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


# Bound per quantity as a multiple of the precision's force tolerance (1e-6 mixed,
# 1e-12 FP64): acc, pot and jerk at the tolerance itself; snap at ten times it (its
# pair term has two more cancelling factors, alpha and beta).
QUANTITY_SCALE = {"acc": 1.0, "pot": 1.0, "jrk": 1.0, "snp": 10.0}


def check_forces(res, ref, order, tol_med, tol_p90=None, label=""):
    """Median (and, with tol_p90, 90th-percentile) per-particle relative errors of every
    computed quantity -- acc, pot, jerk and, for Hermite-6, snap -- as the reference's
    accuracy test prints them (test_gravity_block.cpp:359-362), each asserted against
    its bound (QUANTITY_SCALE x the tolerance).  Returns {quantity: (median, p90, max)}."""
    out = {}
    keys = ["acc", "jrk", "pot"] + (["snp"] if order >= 2 else [])
    for k in keys:
        e = oracle.rel_err(res[k], ref[k])
        out[k] = (float(np.median(e)), float(np.percentile(e, 90)), float(e.max()))
    print(f"{label} rel err median/p90: " + ", ".join(f"{k} {v[0]:.2e}/{v[1]:.2e}" for k, v in out.items()),
          flush=True)
    for k in keys:
        med, p90, _ = out[k]
        bound = tol_med * QUANTITY_SCALE[k]
        assert med <= bound, f"{label} {k} median rel err {med:.3e} > {bound:.1e} ({out})"
        if tol_p90 is not None:
            b90 = tol_p90 * QUANTITY_SCALE[k]
            assert p90 <= b90, f"{label} {k} p90 rel err {p90:.3e} > {b90:.1e} ({out})"
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
