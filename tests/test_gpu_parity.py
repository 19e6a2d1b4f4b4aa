"""GPU parity of the HIP path against the CPU oracle, through the C-ABI.

Bars: nearest neighbour and neighbour lists bit-exact under the working-precision
distance contract (oracle/oracle.c header); predictor bit-exact; forces within the
stated tolerance of the FP64 oracle:
  * mixed precision:  median relative error of acc and pot <= 1e-6
  * all-FP64 kernel:  median relative error of acc and pot <= 1e-12
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
import sapporo2_amd as sa
from sapporo2_amd import ics
from _util import (check_forces, check_nn_lists, load_engine, plummer_system, random_derivs)

pytestmark = pytest.mark.gpu

TOL = {sa.DOUBLESINGLE: 1e-6, sa.DOUBLE: 1e-12}
PRECS = [sa.DOUBLESINGLE, sa.DOUBLE]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", PRECS)
def test_h4_plummer_8192_all_i(prec):
    """Config 1: Plummer N=8192, eps2=1e-4, one all-i x all-j call."""
    _say(f"h4 plummer 8192 prec={prec}")
    n = 8192
    m, x, v, ids, h2 = plummer_system(n, seed=11)
    eps2 = 1e-4
    ref = oracle.force(1, x, v, ids, x, v, m, ids, eps2, h2=h2)
    with sa.Engine(sa.FOURTH, prec) as eng:
        load_engine(eng, m, x, v, ids)
        eng.set_time(0.0)
        eng.start_grav_calc(n, ids, x, v, eps2, h2)
        res = eng.get_grav_results(n, n)
        check_forces(res, ref, 1, TOL[prec], label="plummer8192")
        check_nn_lists(eng, res, prec == sa.DOUBLE, x, ids, x, ids, h2, label="plummer8192")


@pytest.mark.parametrize("prec", PRECS)
def test_h4_reference_recipe_1024(prec):
    """The reference accuracy test's own inputs (test_gravity_block.cpp:40-60, drand48,
    eps2 = 0, h2 = 0.1, ids i+1)."""
    _say(f"h4 drand48 1024 prec={prec}")
    m, x, v, ids, h2 = ics.drand48_block(1024)
    ref = oracle.force(1, x, v, ids, x, v, m, ids, 0.0, h2=h2)
    with sa.Engine(sa.FOURTH, prec) as eng:
        load_engine(eng, m, x, v, ids)
        eng.set_time(0.0)
        eng.start_grav_calc(1024, ids, x, v, 0.0, h2)
        res = eng.get_grav_results(1024, 1024)
        check_forces(res, ref, 1, TOL[prec], label="drand48")
        check_nn_lists(eng, res, prec == sa.DOUBLE, x, ids, x, ids, h2, label="drand48")


@pytest.mark.parametrize("prec", PRECS)
def test_h4_small_active_sets_and_shapes(prec):
    """Active counts 1..1000 at N=8192 with every lane shape P and several j-splits S."""
    n = 8192
    m, x, v, ids, h2 = plummer_system(n, seed=12)
    eps2 = 1e-4
    rng = np.random.Generator(np.random.PCG64(5))
    with sa.Engine(sa.FOURTH, prec) as eng:
        load_engine(eng, m, x, v, ids)
        eng.set_time(0.0)
        for ni in [1, 2, 3, 5, 17, 64, 100, 1000]:
            sel = np.sort(rng.choice(n, ni, replace=False))
            ref = oracle.force(1, x[sel], v[sel], ids[sel], x, v, m, ids, eps2, h2=h2[sel])
            shapes = [(0, 0)] + [(P, S) for P in (1, 4, 16, 64) for S in (1, 7, 64)]
            base = None
            for P, S in shapes:
                _say(f"h4 small ni={ni} P={P} S={S} prec={prec}")
                eng.set_force_override(P, S)
                eng.start_grav_calc(n, ids[sel], x[sel], v[sel], eps2, h2[sel])
                res = eng.get_grav_results(n, ni)
                check_forces(res, ref, 1, TOL[prec] * (10 if ni < 5 else 1), label=f"ni={ni} P={P} S={S}")
                check_nn_lists(eng, res, prec == sa.DOUBLE, x[sel], ids[sel], x, ids, h2[sel],
                               label=f"ni={ni} P={P} S={S}")
                if base is None:
                    base = res
            eng.set_force_override(0, 0)


@pytest.mark.parametrize("prec", PRECS)
def test_predictor_bit_exact_and_block_step(prec):
    """FP64 predictor is bit-exact against the oracle; forces on predicted j match."""
    _say(f"predictor prec={prec}")
    n = 4096
    m, x, v, ids, h2 = plummer_system(n, seed=13)
    acc, jrk, _, _ = random_derivs(n, 3, 0.5)
    rng = np.random.Generator(np.random.PCG64(9))
    tj = -rng.uniform(0.0, 1.0 / 64, n)
    ti = 1.0 / 128
    with sa.Engine(sa.FOURTH, prec) as eng:
        load_engine(eng, m, x, v, ids, acc, jrk, tj=tj)
        eng.set_time(ti)
        sel = np.sort(rng.choice(n, 300, replace=False))
        px, pv, _ = oracle.predict(1, ti, tj, x, v, acc, jrk)
        eng.start_grav_calc(n, ids[sel], px[sel], pv[sel], 1e-4, h2[sel])
        res = eng.get_grav_results(n, len(sel))
        for a in [0, 1, 17, 4095]:
            st = eng.retrieve_j_particle_state(a)
            assert np.array_equal(st["ppos"], px[a]), f"ppos {a}"
            assert np.array_equal(st["pvel"], pv[a]), f"pvel {a}"
            assert np.array_equal(st["pos"], x[a]) and np.array_equal(st["acc"], acc[a])
            assert st["id"] == ids[a] and st["mass"] == m[a]
        ref = oracle.force(1, px[sel], pv[sel], ids[sel], px, pv, m, ids, 1e-4, h2=h2[sel])
        check_forces(res, ref, 1, TOL[prec], label="predicted")
        check_nn_lists(eng, res, prec == sa.DOUBLE, px[sel], ids[sel], px, ids, h2[sel], label="predicted")


@pytest.mark.parametrize("prec", PRECS)
def test_h6_king(prec):
    """Hermite-6 (acc, jerk, snap, pot) on a King W0=6 model with predicted acc."""
    _say(f"h6 king prec={prec}")
    n = 4096
    m, x, v = ics.king(n, 6.0, seed=21)
    ids = np.arange(n, dtype=np.int32)
    acc, jrk, snp, crk = random_derivs(n, 7, 0.3)
    rng = np.random.Generator(np.random.PCG64(10))
    tj = -rng.uniform(0.0, 1.0 / 256, n)
    ti = 1.0 / 512
    eps2 = 1e-6
    with sa.Engine(sa.SIXTH, prec) as eng:
        load_engine(eng, m, x, v, ids, acc, jrk, snp, crk, tj=tj)
        eng.set_time(ti)
        px, pv, pa = oracle.predict(2, ti, tj, x, v, acc, jrk, snp, crk)
        sel = np.sort(rng.choice(n, 1000, replace=False))
        eng.start_grav_calc(n, ids[sel], px[sel], pv[sel], eps2, None, ai=pa[sel])
        res = eng.get_grav_results(n, len(sel))
        ref = oracle.force(2, px[sel], pv[sel], ids[sel], px, pv, m, ids, eps2, ai=pa[sel], aj=pa)
        errs = check_forces(res, ref, 2, TOL[prec], label="h6")
        _say(f"h6 errors {errs}")
        check_nn_lists(eng, res, prec == sa.DOUBLE, px[sel], ids[sel], px, ids, None, check_lists=False)
        p = eng.retrieve_predicted_j_particle(5)
        assert np.array_equal(p["pos"], px[5]) and np.array_equal(p["vel"], pv[5])
        assert np.array_equal(p["acc"], pa[5])
