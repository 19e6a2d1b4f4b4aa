"""GPU parity at the benchmark configurations' own sizes, and exact nearest-neighbour ties.

* The block-step sweep of the metric (Plummer N = 131072, eps2 = 1e-6, t_j on 20 power-of-two
  levels, j-particles predicted to the block time, the previous active set re-sent before each
  block step after the first, so that the particles not yet re-sent are predicted from their own
  t_j, active sets 2^0 .. 2^17 in calls of <= 16384): every bucket is checked on a
  sample of <= 1024 i-particles against the oracle -- forces within tolerance, nearest
  neighbours and neighbour lists bit-exact -- and the device's predicted j-state is checked
  bit-exact against the oracle predictor.
* The 512-cluster ensemble (N log-uniform in [256, 16384], seed 4, eps2 = 1e-4, ~2.2 M
  particles) in one call with uneven active sets: two i-particles of every cluster against
  the oracle, with lists.
* Plummer 2^20 + 1 % central mass (eps2 = 1e-8, t_j on 28 power-of-two levels down to
  2^-30, j-particles predicted to the block time), the whole system active in calls of
  16384: the predicted j-state bit-exact against the oracle predictor, forces of 512 random
  i-particles, the central mass and its 63 nearest stars within tolerance, their nearest
  neighbours and lists bit-exact.
* A lattice with many exact distance ties and ids decreasing with address, over lane shapes
  P in {1, 4, 16, 64} and several j-splits: nearest neighbour and lists bit-exact (ties go to
  the smallest j address, oracle/oracle.c; the reference's first-in-iteration rule,
  kernels.cu:539-546, 994-1003).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
import sapporo2_amd as sa
from sapporo2_amd import ics
from _util import check_forces

pytestmark = pytest.mark.gpu
PRECS = [sa.DOUBLESINGLE, sa.DOUBLE]
TOL = {sa.DOUBLESINGLE: 1e-6, sa.DOUBLE: 1e-12}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _expected_nn(f64, xi, idi, xj, vj, mj, idj, h2, eps2, ref=None):
    if f64:
        r = ref if ref is not None else oracle.force(1, xi, np.zeros_like(xi), idi, xj, vj, mj, idj, eps2, h2=h2)
        return r["nn_addr"], r["nn_ds2"], r["ngb_cnt"], r["ngb_list"]
    r = oracle.neighbours_mixed(xi, idi, xj, idj, h2)
    return r["nn_addr"], r["nn_ds2"], r["ngb_cnt"], r["ngb_list"]


def _compare_nn_lists(got_nnb, got_dsmin, got_lists, nn_addr, nn_ds2, cnt, lst, idj, label):
    nnb, dsmin, found = oracle.api_nn(nn_addr, nn_ds2, idj)
    assert np.array_equal(got_nnb, nnb), f"{label}: nnb mismatches {np.sum(got_nnb != nnb)}"
    assert np.array_equal(np.isnan(got_dsmin), ~found), f"{label}: dsmin written-mask"
    assert np.array_equal(got_dsmin[found], dsmin[found]), f"{label}: dsmin mismatches"
    bad = 0
    for q, (l, n) in enumerate(got_lists):
        if n != cnt[q] or (n < 256 and not np.array_equal(l, lst[q, :n])):
            bad += 1
    assert bad == 0, f"{label}: neighbour-list mismatches {bad}"


@pytest.mark.parametrize("prec", PRECS)
def test_block_step_sweep_n131072(prec):
    n, eps2, seed = 131072, 1e-6, 2024
    _say(f"block-step sweep N={n} prec={prec}")
    m, x, v = ics.plummer(n, seed=seed)
    ids = np.arange(n, dtype=np.int32)
    h2 = ics.neighbour_h2(x, 32.0)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    lev = ics.block_levels(n, 20, seed=seed + 2)
    tj = -rng.uniform(0.0, 1.0, n) * 2.0 ** -(3.0 + lev)
    counts = [2 ** k for k in range(18)]
    blocks = ics.active_subsets(n, counts, seed=seed + 3)
    # j-particles' Taylor coefficients: the oracle's FP64 acc and jerk at t_j
    d = oracle.force_plain(1, x, v, ids, x, v, m, ids, eps2)
    acc, jrk = d["acc"], d["jrk"]
    # the caller's predicted i-particles at the block time t = 0
    px, pv, _ = oracle.predict(1, 0.0, tj, x, v, acc, jrk)
    # j-state as the engine holds it: raw (x, v, t_j), re-sent particles at (px, pv, 0)
    xs, vs, ts = x.copy(), v.copy(), tj.copy()
    smp_rng = np.random.Generator(np.random.PCG64(99))
    with sa.Engine(sa.FOURTH, prec) as eng:
        eng.set_j_particles(np.arange(n), ids, m, x, v, acc, jrk, tj=tj)
        for b, sel in enumerate(blocks):
            if b > 0:
                prev = blocks[b - 1]
                eng.set_j_particles(prev, ids[prev], m[prev], px[prev], pv[prev], acc[prev], jrk[prev], tj=0.0)
                xs[prev], vs[prev], ts[prev] = px[prev], pv[prev], 0.0
            eng.set_time(0.0)
            pj, vj, _ = oracle.predict(1, 0.0, ts, xs, vs, acc, jrk)
            smp = np.sort(smp_rng.choice(len(sel), min(1024, len(sel)), replace=False))
            got = {k: [] for k in ("acc", "jrk", "pot", "nnb", "dsmin")}
            lists = []
            for c0 in range(0, len(sel), sa.NPIPES):
                ch = sel[c0:c0 + sa.NPIPES]
                eng.start_grav_calc(n, ids[ch], px[ch], pv[ch], eps2, h2[ch])
                r = eng.get_grav_results(n, len(ch))
                eng.read_ngb_list()
                loc = smp[(smp >= c0) & (smp < c0 + len(ch))] - c0
                for k in got:
                    got[k].append(r[k][loc])
                lists += [eng.get_ngb_list(int(q))[:2] for q in loc]
            got = {k: np.concatenate(vv) for k, vv in got.items()}
            isel = sel[smp]
            ref = oracle.force(1, px[isel], pv[isel], ids[isel], pj, vj, m, ids, eps2, h2=h2[isel])
            label = f"bucket {len(sel)} prec={prec}"
            check_forces(got, ref, 1, TOL[prec] * (10 if len(isel) < 5 else 1), label=label)
            nn_addr, nn_ds2, cnt, lst = _expected_nn(prec == sa.DOUBLE, px[isel], ids[isel], pj, vj, m, ids,
                                                     h2[isel], eps2, ref)
            _compare_nn_lists(got["nnb"], got["dsmin"], lists, nn_addr, nn_ds2, cnt, lst, ids, label)
            if b in (0, 9, 17):
                probe = np.sort(smp_rng.choice(n, 64, replace=False))
                rp = eng.retrieve_predicted_j_particles(probe)
                assert np.array_equal(rp["pos"], pj[probe]) and np.array_equal(rp["vel"], vj[probe]), label
        _say(f"sweep prec={prec}: {len(blocks)} buckets checked")


@pytest.mark.parametrize("prec", PRECS)
def test_ensemble_512_clusters_full_size(prec):
    _say(f"ensemble 512 clusters prec={prec}")
    sizes = ics.ensemble_sizes(512, 256, 16384, seed=4)
    rng = np.random.Generator(np.random.PCG64(77))
    cl = []
    for c, k in enumerate(sizes):
        m, x, v = ics.plummer(int(k), seed=1000 + c)
        cl.append((m, x, v, ics.neighbour_h2(x, 32.0)))
    eps = np.full(len(sizes), 1e-4)
    with sa.Engine(sa.FOURTH, prec) as eng:
        offs = eng.set_clusters(sizes)
        for c, (m, x, v, _) in enumerate(cl):
            eng.set_j_particles(offs[c] + np.arange(len(m)), np.arange(len(m)), m, x, v)
        eng.set_time(0.0)
        pick = []
        for c, (m, x, v, _) in enumerate(cl):
            k = [1, 2, 17, len(m) // 8, len(m) // 2, len(m)][c % 6]
            pick.append(np.sort(rng.choice(len(m), k, replace=False)))
        cidx = np.concatenate([np.full(len(p), c, np.int32) for c, p in enumerate(pick)])
        X = np.concatenate([cl[c][1][p] for c, p in enumerate(pick)])
        V = np.concatenate([cl[c][2][p] for c, p in enumerate(pick)])
        H = np.concatenate([cl[c][3][p] for c, p in enumerate(pick)])
        I = np.concatenate(pick).astype(np.int32)
        _say(f"ensemble call: {len(I)} i-particles, {int(sizes.sum())} j-particles")
        eng.start_grav_calc_clusters(cidx, I, X, V, eps, H)
        r = eng.get_grav_results(1, len(I))
        # the same results into caller-kept arrays longer than the call, holding stale values
        ni = len(I)
        junk = dict(acc=np.full((ni + 99, 3), 7.0), jrk=np.full((ni + 99, 3), 7.0), pot=np.full(ni + 99, 7.0),
                    nnb=np.full(ni + 99, 7, np.int32), dsmin=np.full(ni + 99, 7.0))
        r2 = eng.get_grav_results(1, ni, out=junk)
        for k in ("acc", "jrk", "pot", "nnb", "dsmin"):
            assert np.array_equal(r2[k], r[k], equal_nan=k == "dsmin"), f"out= rows differ: {k}"
        # keep_dsmin: dsmin written where a nearest neighbour was found, the caller's
        # values (here 5.0) kept elsewhere, as the reference leaves them
        junk["dsmin"].fill(5.0)
        r2 = eng.get_grav_results(1, ni, out=junk, keep_dsmin=True)
        found = ~np.isnan(r["dsmin"])
        assert np.array_equal(r2["dsmin"][found], r["dsmin"][found]), "keep_dsmin: written rows"
        assert np.all(r2["dsmin"][~found] == 5.0) and np.all(junk["dsmin"][ni:] == 5.0), "keep_dsmin: kept rows"
        assert np.array_equal(r2["nnb"], r["nnb"]), "keep_dsmin: nnb"
        # without radii (the single-split items write their own result rows): the same rows
        eng.start_grav_calc_clusters(cidx, I, X, V, eps)
        r3 = eng.get_grav_results(1, ni)
        for k in ("acc", "jrk", "pot", "nnb", "dsmin"):
            assert np.array_equal(r3[k], r[k], equal_nan=k == "dsmin"), f"no-radii rows differ: {k}"
        eng.start_grav_calc_clusters(cidx, I, X, V, eps, H)
        r = eng.get_grav_results(1, ni)
        eng.read_ngb_list()
        starts = np.concatenate([[0], np.cumsum([len(p) for p in pick])[:-1]])
        got = {k: [] for k in ("acc", "jrk", "pot")}
        refs = {k: [] for k in ("acc", "jrk", "pot")}
        for c, p in enumerate(pick):
            m, x, v, h2 = cl[c]
            q = np.arange(min(2, len(p)))
            rows = starts[c] + q
            ii = p[q]
            idc = np.arange(len(m), dtype=np.int32)
            ref = oracle.force(1, x[ii], v[ii], idc[ii], x, v, m, idc, 1e-4, h2=h2[ii])
            for k in got:
                got[k].append(r[k][rows])
                refs[k].append(ref[k])
            nn_addr, nn_ds2, cnt, lst = _expected_nn(prec == sa.DOUBLE, x[ii], idc[ii], x, v, m, idc, h2[ii],
                                                     1e-4, ref)
            lists = [eng.get_ngb_list(int(row))[:2] for row in rows]
            _compare_nn_lists(r["nnb"][rows], r["dsmin"][rows], lists, nn_addr, nn_ds2, cnt, lst, idc,
                              f"cluster {c} prec={prec}")
        check_forces({k: np.concatenate(v) for k, v in got.items()}, {k: np.concatenate(v) for k, v in refs.items()},
                     1, TOL[prec], label=f"512 clusters prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_central_mass_2e20_deep_levels_all_active(prec):
    nstar, eps2 = 1048576, 1e-8
    _say(f"central mass 2^20, 28 levels, all active, prec={prec}")
    m, x, v = ics.plummer_bh(nstar, seed=91, mbh=0.01)
    n = nstar + 1
    ids = np.arange(n, dtype=np.int32)
    h2 = np.concatenate([ics.neighbour_h2(x[:nstar], 32.0, mtot=0.99), [1e-4]])
    rng = np.random.Generator(np.random.PCG64(94))
    # Taylor coefficients of the right scale (the predictor's arithmetic does not depend on
    # their being the system's own derivatives)
    acc = rng.normal(0.0, 1.0, (n, 3))
    jrk = rng.normal(0.0, 10.0, (n, 3))
    lev = ics.block_levels(n, 28, seed=93)
    deep = rng.choice(n, 1024, replace=False)       # the geometric occupancy leaves the
    lev[deep] = 20 + np.arange(1024) % 8            # deepest levels nearly empty: fill them
    tj = -rng.uniform(0.0, 1.0, n) * 2.0 ** -(3.0 + lev)
    pj, vj, _ = oracle.predict(1, 0.0, tj, x, v, acc, jrk)
    rad2 = np.sum(x * x, axis=1)
    smp = np.unique(np.concatenate([rng.choice(n, 512, replace=False), np.argsort(rad2)[:64]]))
    got = {k: np.zeros((len(smp), 3)) for k in ("acc", "jrk")}
    got.update({k: np.zeros(len(smp)) for k in ("pot", "dsmin")})
    got["nnb"] = np.zeros(len(smp), np.int32)
    lists = [None] * len(smp)
    with sa.Engine(sa.FOURTH, prec) as eng:
        eng.set_j_particles(np.arange(n), ids, m, x, v, acc, jrk, tj=tj)
        eng.set_time(0.0)
        for c0 in range(0, n, sa.NPIPES):
            ch = np.arange(c0, min(n, c0 + sa.NPIPES))
            eng.start_grav_calc(n, ids[ch], pj[ch], vj[ch], eps2, h2[ch])
            r = eng.get_grav_results(n, len(ch))
            eng.read_ngb_list()
            pos = np.nonzero((smp >= c0) & (smp < c0 + len(ch)))[0]
            for k in got:
                got[k][pos] = r[k][smp[pos] - c0]
            for q in pos:
                lists[q] = eng.get_ngb_list(int(smp[q] - c0))[:2]
        probe = np.unique(np.concatenate([deep[:256], rng.choice(n, 128, replace=False)]))
        rp = eng.retrieve_predicted_j_particles(probe)
        assert np.array_equal(rp["pos"], pj[probe]) and np.array_equal(rp["vel"], vj[probe]), "predicted j-state"
    _say(f"  {len(smp)} sampled i-particles against the oracle")
    ref = oracle.force(1, pj[smp], vj[smp], ids[smp], pj, vj, m, ids, eps2, h2=h2[smp])
    label = f"central mass 2^20 all active prec={prec}"
    check_forces(got, ref, 1, TOL[prec], label=label)
    nn_addr, nn_ds2, cnt, lst = _expected_nn(prec == sa.DOUBLE, pj[smp], ids[smp], pj, vj, m, ids, h2[smp], eps2,
                                             ref)
    _compare_nn_lists(got["nnb"], got["dsmin"], lists, nn_addr, nn_ds2, cnt, lst, ids, label)


@pytest.mark.parametrize("prec", PRECS)
def test_lattice_ties_decreasing_ids(prec):
    side = 10
    g = (np.arange(side) - side // 2) / 8.0          # exact in float: DS lo parts are 0
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3).astype(np.float64)
    n = len(x)
    v = np.zeros_like(x)
    m = np.full(n, 1.0 / n)
    ids = (n - 1 - np.arange(n)).astype(np.int32)   # ids decrease with address
    eps2 = 1e-4
    # radius 1.5 spacings: the 6 face neighbours (ds2 = 1/64 exactly) plus the 12 edge
    # neighbours (2/64), 18 in the interior; every 7th particle has h2 = 1/64 exactly (on the
    # strict boundary: no neighbour counts)
    h2 = np.full(n, 2.25 / 64.0)
    h2[::7] = 1.0 / 64.0
    rng = np.random.Generator(np.random.PCG64(8))
    with sa.Engine(sa.FOURTH, prec) as eng:
        eng.set_j_particles(np.arange(n), ids, m, x, v)
        eng.set_time(0.0)
        for ni in (1, 3, 40, 200, n):
            sel = np.sort(rng.choice(n, ni, replace=False))
            nn_addr, nn_ds2, cnt, lst = _expected_nn(prec == sa.DOUBLE, x[sel], ids[sel], x, v, m, ids, h2[sel],
                                                     eps2)
            ties = np.sum(np.isclose(nn_ds2, 1.0 / 64.0))
            for P in (1, 4, 16, 64):
                for S in (1, 3, 8, 40):
                    eng.set_force_override(P, S)
                    eng.start_grav_calc(n, ids[sel], x[sel], v[sel], eps2, h2[sel])
                    r = eng.get_grav_results(n, ni)
                    eng.read_ngb_list()
                    lists = [eng.get_ngb_list(q)[:2] for q in range(ni)]
                    _compare_nn_lists(r["nnb"], r["dsmin"], lists, nn_addr, nn_ds2, cnt, lst, ids,
                                      f"lattice ni={ni} P={P} S={S} prec={prec}")
            eng.set_force_override(0, 0)
            _say(f"lattice ni={ni} prec={prec}: {ties} tied minima, shapes P x S agree with the oracle")
