"""Array-form read-back of predicted j-particles (rows gathered on the device for a
list of addresses) against the per-particle read-back, which copies the whole
store back as the reference does (sapporohostclass.cpp:969-1041), and against the
oracle's predictor: unsorted addresses with repeats, before and after updates that
have not been followed by a force call, an empty list, an address never set."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
import sapporo2_amd as sa
from _util import load_engine, plummer_system, random_derivs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _per_particle(eng, addrs, order):
    rows = [eng.retrieve_predicted_j_particle(int(a)) for a in addrs]
    out = dict(mass=np.array([r["mass"] for r in rows]), pos=np.array([r["pos"] for r in rows]),
               vel=np.array([r["vel"] for r in rows]))
    out["acc"] = np.array([r["acc"] for r in rows]) if order >= sa.SIXTH else np.zeros((len(addrs), 3))
    return out


@pytest.mark.parametrize("order", [sa.FOURTH, sa.SIXTH])
@pytest.mark.parametrize("prec", [sa.DOUBLESINGLE, sa.DOUBLE])
def test_array_retrieve(order, prec):
    n = 5000
    m, x, v, ids, h2 = plummer_system(n, seed=71)
    acc, jrk, snp, crk = random_derivs(n, 72, 0.3)
    if order < sa.SIXTH:
        snp = crk = None
    rng = np.random.Generator(np.random.PCG64(73))
    tj = -rng.uniform(0, 1.0 / 256, n)
    ti = 1.0 / 512
    with sa.Engine(order, prec) as eng:
        load_engine(eng, m, x, v, ids, acc, jrk, snp, crk, tj=tj)
        eng.set_time(ti)
        eng.start_grav_calc(n, ids[:64], x[:64], v[:64], 1e-4, None,
                            ai=acc[:64] if order >= sa.SIXTH else None)
        eng.get_grav_results(n, 64)
        addrs = rng.integers(0, n, 777)
        addrs[:5] = [n - 1, 0, 17, 17, n - 1]
        _say(f"array retrieve order={order} prec={prec} n={len(addrs)}")
        got = eng.retrieve_predicted_j_particles(addrs)
        ref = _per_particle(eng, addrs, order)
        for k in ("mass", "pos", "vel", "acc"):
            assert np.array_equal(got[k], ref[k]), k
        px, pv, pa = oracle.predict(order, ti, tj, x, v, acc, jrk, snp, crk)
        assert np.array_equal(got["pos"], px[addrs]) and np.array_equal(got["vel"], pv[addrs])
        if order >= sa.SIXTH:
            assert np.array_equal(got["acc"], pa[addrs])
        assert np.array_equal(got["mass"], m[addrs])

        # updates not yet followed by a force call are flushed before the gather
        upd = np.array([3, 17, 4000])
        m2 = m.copy(); m2[upd] = [0.5, 0.25, 0.125]
        eng.set_j_particles(upd, ids[upd], m2[upd], x[upd] + 0.01, v[upd], None if acc is None else acc[upd],
                            None if jrk is None else jrk[upd], None if snp is None else snp[upd],
                            None if crk is None else crk[upd], tj=tj[upd])
        sel = np.array([17, 3, 4000, 9, 3])
        got = eng.retrieve_predicted_j_particles(sel)
        ref = _per_particle(eng, sel, order)
        for k in ("mass", "pos", "vel", "acc"):
            assert np.array_equal(got[k], ref[k]), f"after update: {k}"
        assert np.array_equal(got["mass"], m2[sel])

        empty = eng.retrieve_predicted_j_particles(np.zeros(0, np.int32))
        assert empty["pos"].shape == (0, 3)
        with pytest.raises(RuntimeError):
            eng.retrieve_predicted_j_particles(np.array([0, n + 10]))
