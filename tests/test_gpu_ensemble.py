"""Ensembles of independent clusters in one launch (the batched-ensemble workload):
every cluster's forces, nearest neighbours and lists must equal those of the same
cluster evaluated alone by the oracle."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
import sapporo2_amd as sa
from sapporo2_amd import ics
from _util import check_forces

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", [sa.DOUBLESINGLE, sa.DOUBLE])
@pytest.mark.parametrize("local_ids", [True, False])
def test_ensemble_matches_clusters_alone(prec, local_ids):
    _say(f"ensemble prec={prec} local_ids={local_ids}")
    rng = np.random.Generator(np.random.PCG64(71))
    sizes = ics.ensemble_sizes(24, 256, 4096, seed=72)
    eps2 = rng.choice([1e-4, 1e-3], len(sizes))
    cl = []
    for c, n in enumerate(sizes):
        m, x, v = ics.plummer(int(n), seed=100 + c)
        ids = np.arange(n, dtype=np.int32) if local_ids else (np.arange(n) + 100000 * c).astype(np.int32)
        cl.append((m, x, v, ids, ics.neighbour_h2(x, 24.0)))
    with sa.Engine(sa.FOURTH, prec) as eng:
        offs = eng.set_clusters(sizes)
        for c, (m, x, v, ids, h2) in enumerate(cl):
            eng.set_j_particles(offs[c] + np.arange(len(m)), ids, m, x, v)
        eng.set_time(0.0)
        # uneven active sets: none, one, a few, many, all
        pick = []
        for c, (m, x, v, ids, h2) in enumerate(cl):
            n = len(m)
            k = [0, 1, 3, 17, n // 3, n][c % 6]
            pick.append(np.sort(rng.choice(n, k, replace=False)))
        cidx = np.concatenate([np.full(len(p), c, np.int32) for c, p in enumerate(pick)])
        X = np.concatenate([cl[c][1][p] for c, p in enumerate(pick)])
        V = np.concatenate([cl[c][2][p] for c, p in enumerate(pick)])
        I = np.concatenate([cl[c][3][p] for c, p in enumerate(pick)])
        H = np.concatenate([cl[c][4][p] for c, p in enumerate(pick)])
        eng.start_grav_calc_clusters(cidx, I, X, V, eps2, H)
        r = eng.get_grav_results(1, len(I))
        eng.read_ngb_list()
        o = 0
        tol = 1e-6 if prec == sa.DOUBLESINGLE else 1e-12
        for c, p in enumerate(pick):
            k = len(p)
            if k == 0:
                continue
            m, x, v, ids, h2 = cl[c]
            sl = slice(o, o + k)
            ref = oracle.force(1, x[p], v[p], ids[p], x, v, m, ids, eps2[c], h2=h2[p])
            sub = {kk: r[kk][sl] for kk in ("acc", "jrk", "pot")}
            check_forces(sub, ref, 1, tol * (10 if k < 5 else 1), label=f"cluster {c}")
            if prec == sa.DOUBLESINGLE:
                nn = oracle.neighbours_mixed(x[p], ids[p], x, ids, h2[p])
            else:
                nn = ref
            nnb, dsmin, found = oracle.api_nn(nn["nn_addr"], nn["nn_ds2"], ids)
            assert np.array_equal(r["nnb"][sl], nnb), f"cluster {c} nnb"
            assert np.array_equal(r["dsmin"][sl][found], dsmin[found]), f"cluster {c} dsmin"
            for q in range(k):
                got, cnt, _ = eng.get_ngb_list(o + q)
                assert cnt == nn["ngb_cnt"][q], (c, q)
                if cnt < 256:
                    assert np.array_equal(got, nn["ngb_list"][q, :cnt]), (c, q)
            o += k
        assert o == len(I)
        # without radii the single-split items write their result rows themselves (the
        # finalize skips them): bitwise the rows of the call with radii
        eng.start_grav_calc_clusters(cidx, I, X, V, eps2)
        r0 = eng.get_grav_results(1, len(I))
        for k in ("acc", "jrk", "pot", "nnb", "dsmin"):
            assert np.array_equal(r0[k], r[k], equal_nan=k == "dsmin"), f"no-radii rows differ: {k}"
        # i-particles must be grouped by cluster
        bad = cidx.copy()
        if len(bad) > 2:
            bad[[0, -1]] = bad[[-1, 0]]
            with pytest.raises(RuntimeError):
                eng.start_grav_calc_clusters(bad, I, X, V, eps2, H)


@pytest.mark.parametrize("prec", [sa.DOUBLESINGLE, sa.DOUBLE])
def test_ensemble_h6_predicted_no_lists(prec):
    """Hermite-6 ensemble on predicted j-particles without neighbour lists: snap and the
    nearest neighbour per cluster against the oracle on the predicted state."""
    _say(f"ensemble h6 prec={prec}")
    rng = np.random.Generator(np.random.PCG64(75))
    sizes = ics.ensemble_sizes(12, 256, 2048, seed=76)
    eps2 = rng.choice([1e-4, 1e-3], len(sizes))
    ti = 1.0 / 1024
    cl = []
    for c, n in enumerate(sizes):
        m, x, v = ics.plummer(int(n), seed=200 + c)
        ids = np.arange(n, dtype=np.int32)
        acc, jrk, snp, crk = (rng.normal(0, 0.3, (n, 3)) for _ in range(4))
        tj = -rng.uniform(0, 1.0 / 512, n)
        cl.append((m, x, v, ids, acc, jrk, snp, crk, tj))
    with sa.Engine(sa.SIXTH, prec) as eng:
        offs = eng.set_clusters(sizes)
        for c, (m, x, v, ids, acc, jrk, snp, crk, tj) in enumerate(cl):
            eng.set_j_particles(offs[c] + np.arange(len(m)), ids, m, x, v, acc, jrk, snp, crk, tj=tj)
        eng.set_time(ti)
        pick, pred = [], []
        for c, (m, x, v, ids, acc, jrk, snp, crk, tj) in enumerate(cl):
            n = len(m)
            k = [1, 5, n // 2, n][c % 4]
            pick.append(np.sort(rng.choice(n, k, replace=False)))
            pred.append(oracle.predict(2, ti, tj, x, v, acc, jrk, snp, crk))
        cidx = np.concatenate([np.full(len(p), c, np.int32) for c, p in enumerate(pick)])
        X = np.concatenate([pred[c][0][p] for c, p in enumerate(pick)])
        V = np.concatenate([pred[c][1][p] for c, p in enumerate(pick)])
        A = np.concatenate([pred[c][2][p] for c, p in enumerate(pick)])
        I = np.concatenate([cl[c][3][p] for c, p in enumerate(pick)])
        eng.start_grav_calc_clusters(cidx, I, X, V, eps2, None, ai=A)
        r = eng.get_grav_results(1, len(I))
        o = 0
        tol = 1e-6 if prec == sa.DOUBLESINGLE else 1e-12
        for c, p in enumerate(pick):
            k = len(p)
            m, ids = cl[c][0], cl[c][3]
            px, pv, pa = pred[c]
            sl = slice(o, o + k)
            ref = oracle.force(2, px[p], pv[p], ids[p], px, pv, m, ids, eps2[c], ai=pa[p], aj=pa)
            sub = {kk: r[kk][sl] for kk in ("acc", "jrk", "snp", "pot")}
            errs = check_forces(sub, ref, 2, tol * (10 if k < 5 else 1), label=f"h6 cluster {c}")
            if prec == sa.DOUBLE:
                assert errs["snp"][0] < 1e-10, (c, errs["snp"])
            nn = ref if prec == sa.DOUBLE else oracle.neighbours_mixed(px[p], ids[p], px, ids, np.zeros(k))
            nnb, dsmin, found = oracle.api_nn(nn["nn_addr"], nn["nn_ds2"], ids)
            assert np.array_equal(r["nnb"][sl], nnb), f"h6 cluster {c} nnb"
            assert np.array_equal(r["dsmin"][sl][found], dsmin[found]), f"h6 cluster {c} dsmin"
            o += k
        assert o == len(I)
