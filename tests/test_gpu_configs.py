"""The larger configurations at their real sizes:

* Hermite-6 on a King W0 = 6 model, N = 262144, eps2 = 1e-6: mixed and FP64 kernels,
  per-quantity errors (acc, jerk, snap, pot) against the FP64 oracle on a sample.
* Plummer N = 1048576 with a central point mass of 1% at rest, stars eps2 = 1e-8:
  active sets of 1024 and 32768; nearest neighbours and lists bit-exact near the
  central mass (dynamic range), forces within tolerance.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
import sapporo2_amd as sa
from sapporo2_amd import ics
from _util import check_forces, check_nn_lists

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", [sa.DOUBLESINGLE, sa.DOUBLE])
def test_h6_king_262144(prec):
    _say(f"h6 king 262144 prec={prec}")
    n = 262144
    m, x, v = ics.king(n, 6.0, seed=81)
    ids = np.arange(n, dtype=np.int32)
    rng = np.random.Generator(np.random.PCG64(81))
    acc, jrk, snp, crk = (rng.normal(0, 0.3, (n, 3)) for _ in range(4))
    tj = -rng.uniform(0, 1.0 / 1024, n)
    ti = 1.0 / 2048
    eps2 = 1e-6
    px, pv, pa = oracle.predict(2, ti, tj, x, v, acc, jrk, snp, crk)
    sel = np.sort(rng.choice(n, 16384, replace=False))
    with sa.Engine(sa.SIXTH, prec) as eng:
        eng.set_j_particles(np.arange(n), ids, m, x, v, acc, jrk, snp, crk, tj=tj)
        eng.set_time(ti)
        eng.start_grav_calc(n, ids[sel], px[sel], pv[sel], eps2, None, ai=pa[sel])
        r = eng.get_grav_results(n, len(sel))
        sub = np.arange(0, len(sel), 32)
        ref = oracle.force(2, px[sel][sub], pv[sel][sub], ids[sel][sub], px, pv, m, ids, eps2,
                           ai=pa[sel][sub], aj=pa)
        rs = {k: (None if r[k] is None else r[k][sub]) for k in r}
        errs = check_forces(rs, ref, 2, 1e-6 if prec == sa.DOUBLESINGLE else 1e-12, label="king")
        _say(f"king H6 prec={prec} median/p90/max rel errors: {errs}")
        if prec == sa.DOUBLE:
            for k in ("jrk", "snp"):
                assert errs[k][0] < 1e-11, (k, errs[k])
        else:
            assert errs["jrk"][0] < 1e-5 and errs["snp"][0] < 1e-4, errs
        check_nn_lists(eng, rs, prec == sa.DOUBLE, px[sel][sub], ids[sel][sub], px, ids, None,
                       check_lists=False, label="king")


@pytest.mark.parametrize("prec", [sa.DOUBLESINGLE, sa.DOUBLE])
def test_plummer_1m_central_mass(prec):
    _say(f"plummer 1M + BH prec={prec}")
    nstar = 1048576
    m, x, v = ics.plummer_bh(nstar, seed=91, mbh=0.01)
    n = nstar + 1
    ids = np.arange(n, dtype=np.int32)
    h2 = np.concatenate([ics.neighbour_h2(x[:nstar], 32.0, mtot=0.99), [1e-4]])
    eps2 = 1e-8
    rad = np.sqrt((x * x).sum(1))
    near = np.argsort(rad)[:256]                   # the BH (radius 0) and its closest stars
    rng = np.random.Generator(np.random.PCG64(92))
    with sa.Engine(sa.FOURTH, prec) as eng:
        eng.set_j_particles(np.arange(n), ids, m, x, v)
        eng.set_time(0.0)
        for k, sel in ((1024, np.sort(rng.choice(n, 1024, replace=False))), (256, np.sort(near)),
                       (32768, np.sort(rng.choice(n, 32768, replace=False)))):
            _say(f"  active set {len(sel)}")
            r = eng.calc(n, ids[sel], x[sel], v[sel], eps2, h2[sel])
            sub = np.arange(len(sel)) if len(sel) <= 1024 else np.arange(0, len(sel), 64)
            ref = oracle.force(1, x[sel][sub], v[sel][sub], ids[sel][sub], x, v, m, ids, eps2, h2=h2[sel][sub])
            rs = {kk: (None if r[kk] is None else r[kk][sub]) for kk in r}
            check_forces(rs, ref, 1, 1e-6 if prec == sa.DOUBLESINGLE else 1e-12, label=f"bh {len(sel)}")
            if len(sel) <= 16384:   # lists of the last call only
                check_nn_lists(eng, rs if len(sub) == len(sel) else r, prec == sa.DOUBLE, x[sel], ids[sel], x, ids,
                               h2[sel], label=f"bh {len(sel)}")
        # the central mass is every near star's nearest neighbour candidate: its own NN and
        # the stars' NN ids must match the oracle exactly (checked above for the 256 nearest)
