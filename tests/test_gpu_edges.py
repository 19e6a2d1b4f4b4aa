"""Edge cases of the force path against the oracle:

* nj smaller than the store and not a multiple of the 64-address tile (the API's
  nj = number of addresses [0, nj) used, sapporohostclass.cpp:718);
* coincident particles (ds2 = 0) and duplicate ids (self-exclusion is by id,
  kernels.cu:565, so every j sharing the i-particle's id is skipped);
* an address re-set twice before the next call (the last values win,
  sapporohostclass.cpp:482-490);
* get_ngb_list with maxlength at or below the count (overflow flag, full count,
  a sorted subset of the list; sapporohostclass.cpp:837-880).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
import sapporo2_amd as sa
from _util import check_forces, check_nn_lists, load_engine, plummer_system

pytestmark = pytest.mark.gpu
PRECS = [sa.DOUBLESINGLE, sa.DOUBLE]
TOL = {sa.DOUBLESINGLE: 1e-6, sa.DOUBLE: 1e-12}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", PRECS)
def test_partial_nj(prec):
    n = 4096
    m, x, v, ids, h2 = plummer_system(n, seed=41)
    eps2 = 1e-4
    rng = np.random.Generator(np.random.PCG64(41))
    sel = np.sort(rng.choice(np.arange(2048, n), 100, replace=False))   # ids outside [0, 2048)
    with sa.Engine(sa.FOURTH, prec) as eng:
        load_engine(eng, m, x, v, ids)
        eng.set_time(0.0)
        for nj in (1, 63, 64, 65, 1000, 2047):
            _say(f"partial nj={nj} prec={prec}")
            eng.start_grav_calc(nj, ids[sel], x[sel], v[sel], eps2, h2[sel])
            res = eng.get_grav_results(nj, len(sel))
            ref = oracle.force(1, x[sel], v[sel], ids[sel], x[:nj], v[:nj], m[:nj], ids[:nj], eps2, h2=h2[sel])
            check_forces(res, ref, 1, TOL[prec], label=f"nj={nj}")
            check_nn_lists(eng, res, prec == sa.DOUBLE, x[sel], ids[sel], x[:nj], ids[:nj], h2[sel],
                           label=f"nj={nj}")


@pytest.mark.parametrize("prec", PRECS)
def test_coincident_and_duplicate_ids(prec):
    n = 512
    m, x, v, ids, h2 = plummer_system(n, seed=42)
    x = x.copy(); ids = ids.copy()
    x[10] = x[11]                # coincident, distinct ids: ds2 = 0 exactly
    x[300] = x[301] = x[302]     # three at one point
    ids[21] = ids[20]            # duplicate id at distinct positions: mutually excluded
    ids[401] = ids[400] = ids[399]
    eps2 = 1e-4
    with sa.Engine(sa.FOURTH, prec) as eng:
        load_engine(eng, m, x, v, ids)
        eng.set_time(0.0)
        _say(f"coincident/duplicate prec={prec}")
        eng.start_grav_calc(n, ids, x, v, eps2, h2)
        res = eng.get_grav_results(n, n)
        ref = oracle.force(1, x, v, ids, x, v, m, ids, eps2, h2=h2)
        assert np.all(np.isfinite(res["acc"])) and np.all(np.isfinite(res["pot"]))
        check_forces(res, ref, 1, TOL[prec], label="coincident")
        check_nn_lists(eng, res, prec == sa.DOUBLE, x, ids, x, ids, h2, label="coincident")
        assert res["dsmin"][10] == 0.0 and res["nnb"][10] == ids[11]
        assert res["nnb"][300] == ids[301] and res["dsmin"][300] == 0.0   # tie at 0: smallest address
        l20, c20, _ = eng.get_ngb_list(20)
        assert ids[20] not in l20


@pytest.mark.parametrize("prec", PRECS)
def test_overwrite_before_calc(prec):
    n = 1024
    m, x, v, ids, h2 = plummer_system(n, seed=43)
    x = x.copy(); m = m.copy()
    with sa.Engine(sa.FOURTH, prec) as eng:
        load_engine(eng, m, x, v, ids)
        eng.set_time(0.0)
        eng.start_grav_calc(n, ids[:8], x[:8], v[:8], 1e-4, h2[:8])
        eng.get_grav_results(n, 8)
        # address 5 set twice, address 700 once, before the next call
        eng.set_j_particles(np.array([5]), ids[[5]], np.array([0.5]), np.array([[0.1, 0.2, 0.3]]),
                            np.zeros((1, 3)))
        m[5] = 0.25; x[5] = [0.3, -0.2, 0.1]
        m[700] = 0.125; x[700] = [-0.4, 0.4, 0.0]
        eng.set_j_particles(np.array([5, 700]), ids[[5, 700]], m[[5, 700]], x[[5, 700]], v[[5, 700]])
        eng.set_time(0.0)
        sel = np.arange(0, n, 7)
        _say(f"overwrite prec={prec}")
        eng.start_grav_calc(n, ids[sel], x[sel], v[sel], 1e-4, h2[sel])
        res = eng.get_grav_results(n, len(sel))
        ref = oracle.force(1, x[sel], v[sel], ids[sel], x, v, m, ids, 1e-4, h2=h2[sel])
        check_forces(res, ref, 1, TOL[prec], label="overwrite")
        check_nn_lists(eng, res, prec == sa.DOUBLE, x[sel], ids[sel], x, ids, h2[sel], label="overwrite")
        st = eng.retrieve_j_particle_state(5)
        assert st["mass"] == 0.25 and np.array_equal(st["pos"], x[5])


@pytest.mark.parametrize("prec", PRECS)
def test_get_list_maxlength(prec):
    n = 2048
    m, x, v, ids, h2 = plummer_system(n, seed=44, mean_ngb=64.0)
    with sa.Engine(sa.FOURTH, prec) as eng:
        load_engine(eng, m, x, v, ids)
        eng.set_time(0.0)
        sel = np.arange(0, n, 16)
        _say(f"maxlength prec={prec}")
        eng.start_grav_calc(n, ids[sel], x[sel], v[sel], 1e-4, h2[sel])
        eng.get_grav_results(n, len(sel))
        if prec == sa.DOUBLE:
            r = oracle.force(1, x[sel], np.zeros_like(x[sel]), ids[sel], x, np.zeros_like(x), np.ones(n), ids, 1.0,
                             h2=h2[sel])
        else:
            r = oracle.neighbours_mixed(x[sel], ids[sel], x, ids, h2[sel])
        eng.read_ngb_list()
        tested = 0
        for i in range(len(sel)):
            c = int(r["ngb_cnt"][i])
            if not 4 <= c < 256:
                continue
            full = r["ngb_list"][i, :c]
            for ml in (c + 1, c, c - 3):
                l, cnt, ov = eng.get_ngb_list(i, ml)
                assert cnt == c, (i, ml, cnt, c)
                assert ov == int(c >= ml), (i, ml, ov)
                k = min(c, ml)
                assert len(l) == k and np.all(np.diff(l) > 0) and np.all(np.isin(l, full)), (i, ml)
                if ml > c:
                    assert np.array_equal(l, full)
            tested += 1
        assert tested > 10
