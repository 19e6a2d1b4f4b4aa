"""Checkpoint of the j-store: a dump loaded into a fresh engine (which held other
particles) must reproduce the source engine bit for bit -- the next force calls,
neighbour lists and read-backs, and a second dump of the loaded engine equals the
first.  Covered: dumps with staged updates (flushed into the dump), with a
prediction pending (set_time without a force call yet), an ensemble's cluster
layout, and the refusals (other precision, truncated buffer, not a dump)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import sapporo2_amd as sa
from _util import load_engine, plummer_system, random_derivs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _call(eng, n, ids, x, v, eps2, h2, ai):
    eng.start_grav_calc(n, ids, x, v, eps2, h2, ai=ai)
    r = eng.get_grav_results(n, len(ids))
    lists = None
    if h2 is not None:
        eng.read_ngb_list()
        lists = [eng.get_ngb_list(i)[:2] for i in range(len(ids))]
    return r, lists


def _same(a, b, label):
    (ra, la), (rb, lb) = a, b
    for k in ("acc", "jrk", "snp", "pot", "nnb"):
        if ra.get(k) is not None:
            assert np.array_equal(ra[k], rb[k]), f"{label}: {k}"
    assert np.array_equal(ra["dsmin"], rb["dsmin"], equal_nan=True), f"{label}: dsmin"
    if la is not None:
        for i, ((l1, n1), (l2, n2)) in enumerate(zip(la, lb)):
            assert n1 == n2 and (n1 >= 256 or np.array_equal(l1, l2)), f"{label}: list {i}"


@pytest.mark.parametrize("order", [sa.FOURTH, sa.SIXTH])
@pytest.mark.parametrize("prec", [sa.DOUBLESINGLE, sa.DOUBLE])
def test_checkpoint_roundtrip(order, prec):
    n = 6000
    m, x, v, ids, h2 = plummer_system(n, seed=81)
    acc, jrk, snp, crk = random_derivs(n, 82, 0.3)
    if order < sa.SIXTH:
        snp = crk = None
    rng = np.random.Generator(np.random.PCG64(83))
    tj = -rng.uniform(0, 1.0 / 256, n)
    eps2 = 1e-4
    six = order >= sa.SIXTH
    src, dst = sa.Engine(order, prec), sa.Engine(order, prec)
    try:
        load_engine(src, m, x, v, ids, acc, jrk, snp, crk, tj=tj)
        # the destination holds a different, larger system before the load
        m2, x2, v2, ids2, _ = plummer_system(n + 3000, seed=84)
        load_engine(dst, m2, x2, v2, ids2)
        dst.set_time(0.0)
        dst.start_grav_calc(n + 3000, ids2[:10], x2[:10], v2[:10], eps2)
        dst.get_grav_results(n + 3000, 10)

        src.set_time(1.0 / 512)
        sel = np.sort(rng.choice(n, 500, replace=False))
        _call(src, n, ids[sel], x[sel], v[sel], eps2, h2[sel], acc[sel] if six else None)
        upd = np.array([4, 77, 5999])
        src.set_j_particles(upd, ids[upd], m[upd] * 2, x[upd] + 0.01, v[upd], acc[upd], jrk[upd],
                            None if snp is None else snp[upd], None if crk is None else crk[upd],
                            tj=tj[upd] + 1.0 / 1024)
        for pending in (False, True):
            if pending:
                src.set_time(3.0 / 1024)     # dump with the prediction still to run
            _say(f"checkpoint order={order} prec={prec} pending={pending}")
            dump = src.save_state()
            dst.load_state(dump)
            assert np.array_equal(dst.save_state(), src.save_state()), "dump of the loaded engine differs"
            for c, ti in enumerate((None, 5.0 / 1024)):
                if ti is not None:
                    for e in (src, dst):
                        e.set_time(ti)
                sel = np.sort(rng.choice(n, 300 + 100 * c, replace=False))
                out = [_call(e, n, ids[sel], x[sel], v[sel], eps2, h2[sel], acc[sel] if six else None)
                       for e in (src, dst)]
                _same(out[0], out[1], f"pending={pending} call {c}")
            addrs = np.array([0, 4, 77, 5999, 1234])
            a, b = src.retrieve_predicted_j_particles(addrs), dst.retrieve_predicted_j_particles(addrs)
            for k in a:
                assert np.array_equal(a[k], b[k]), k
            sa_, sb = src.retrieve_j_particle_state(77), dst.retrieve_j_particle_state(77)
            for k in sa_:
                assert np.array_equal(np.asarray(sa_[k]), np.asarray(sb[k])), k

        other = sa.Engine(order, sa.DOUBLE if prec == sa.DOUBLESINGLE else sa.DOUBLESINGLE)
        try:
            with pytest.raises(RuntimeError):
                other.load_state(dump)
        finally:
            other.close()
        with pytest.raises(RuntimeError):
            dst.load_state(dump[: len(dump) // 2])
        with pytest.raises(RuntimeError):
            dst.load_state(np.zeros(256, np.uint8))
    finally:
        src.close()
        dst.close()


def test_checkpoint_clusters():
    sizes = [300, 1000, 64, 777]
    src, dst = sa.Engine(sa.FOURTH, sa.DOUBLESINGLE), sa.Engine(sa.FOURTH, sa.DOUBLESINGLE)
    try:
        offs = src.set_clusters(sizes)
        rng = np.random.Generator(np.random.PCG64(85))
        cl, I, X, V, H = [], [], [], [], []
        for c, (nc, off) in enumerate(zip(sizes, offs)):
            m, x, v, ids, h2 = plummer_system(nc, seed=90 + c)
            src.set_j_particles(off + np.arange(nc), ids, m, x, v)
            k = np.sort(rng.choice(nc, min(nc, 100), replace=False))
            cl.append(np.full(len(k), c)); I.append(ids[k]); X.append(x[k]); V.append(v[k]); H.append(h2[k])
        cl, I, X, V, H = (np.concatenate(a) for a in (cl, I, X, V, H))
        eps2 = np.array([1e-4, 1e-3, 1e-5, 1e-4])
        src.set_time(0.0)
        _say("checkpoint clusters")
        dst.load_state(src.save_state())
        out = []
        for e in (src, dst):
            e.start_grav_calc_clusters(cl, I, X, V, eps2, H)
            r = e.get_grav_results(1, len(I))
            e.read_ngb_list()
            out.append((r, [e.get_ngb_list(i)[:2] for i in range(len(I))]))
        _same(out[0], out[1], "clusters")
        assert np.array_equal(dst.save_state(), src.save_state())
    finally:
        src.close()
        dst.close()
