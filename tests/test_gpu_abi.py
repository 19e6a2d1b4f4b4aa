"""GPU tests through the drop-in C-ABIs themselves (GRAPE-6 g6_* in libsapporo_amd.so,
6th-order in libsapporo6th_amd.so), their error paths, the external integrator programs,
and size-independent properties at the full N = 131072."""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
import sapporo2_amd as sa
from sapporo2_amd import ics
from _util import plummer_system

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

c_int_p = ctypes.POINTER(ctypes.c_int)
c_dbl_p = ctypes.POINTER(ctypes.c_double)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def P(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


def I(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def D(v):
    return ctypes.byref(ctypes.c_double(float(v)))


@pytest.fixture(scope="module")
def g6():
    L = ctypes.CDLL(sa.LIB_PATH)
    L.g6_get_neighbour_list_.restype = ctypes.c_int
    return L


def _g6_load(L, m, x, v, ids, a2, j6, tj):
    k18 = np.zeros(3)
    for i in range(len(m)):
        L.g6_set_j_particle_(I(0), I(i), I(ids[i]), D(tj[i]), D(tj[i]), D(m[i]), P(k18),
                             P(np.ascontiguousarray(j6[i])), P(np.ascontiguousarray(a2[i])),
                             P(np.ascontiguousarray(v[i])), P(np.ascontiguousarray(x[i])))


def test_g6_abi_end_to_end(g6):
    """g6_open/set_j_particle (a/2, j/6 inputs)/set_ti/firsthalf/lasthalf2/lists."""
    _say("g6 abi end to end")
    L = g6
    n = 2048
    m, x, v, ids, h2 = plummer_system(n, seed=31)
    rng = np.random.Generator(np.random.PCG64(31))
    acc = rng.normal(0, 0.5, (n, 3)); jrk = rng.normal(0, 0.5, (n, 3))
    tj = -rng.uniform(0, 1.0 / 32, n)
    assert L.g6_open_special(1, (ctypes.c_int * 1)(0)) == 0
    try:
        assert L.g6_npipes_() == 16384
        _g6_load(L, m, x, v, ids, acc / 2, jrk / 6, tj)
        ti = 1.0 / 64
        L.g6_set_ti_(I(0), D(ti))
        px, pv, _ = oracle.predict(1, ti, tj, x, v, acc, jrk)
        sel = np.sort(rng.choice(n, 700, replace=False))
        ni = len(sel)
        xi = np.ascontiguousarray(px[sel]); vi = np.ascontiguousarray(pv[sel])
        idi = np.ascontiguousarray(ids[sel]); h2i = np.ascontiguousarray(h2[sel])
        eps2 = 1e-4
        A = np.zeros((ni, 3)); J = np.zeros((ni, 3)); pot = np.zeros(ni); inn = np.zeros(ni, np.int32)
        L.g6calc_firsthalf_(I(0), I(n), I(ni), P(idi, ctypes.c_int), P(xi), P(vi), P(A), P(J), P(pot),
                            D(eps2), P(h2i))
        rc = L.g6calc_lasthalf2_(I(0), I(n), I(ni), P(idi, ctypes.c_int), P(xi), P(vi), D(eps2), P(h2i),
                                 P(A), P(J), P(pot), P(inn, ctypes.c_int))
        assert rc == 0
        ref = oracle.force(1, xi, vi, idi, px, pv, m, ids, eps2, h2=h2i)
        assert np.median(oracle.rel_err(A, ref["acc"])) < 1e-6
        assert np.median(oracle.rel_err(pot, ref["pot"])) < 1e-6
        mx = oracle.neighbours_mixed(xi, idi, px, ids, h2i)
        nnb, _, _ = oracle.api_nn(mx["nn_addr"], mx["nn_ds2"], ids)
        assert np.array_equal(inn, nnb)
        # lists through the g6 calls, including the maxlength rule (overflow iff count >= maxlength)
        # some outskirts particles have >= 256 neighbours: the read reports the overflow
        assert L.g6_read_neighbour_list_(I(0)) == int(np.any(mx["ngb_cnt"] >= 256))
        buf = np.zeros(256, np.int32)
        nn = ctypes.c_int(0)
        for k in range(ni):
            ov = L.g6_get_neighbour_list_(I(0), I(k), I(256), ctypes.byref(nn), P(buf, ctypes.c_int))
            c = mx["ngb_cnt"][k]
            assert nn.value == c
            if c < 256:
                assert ov == 0 and np.array_equal(buf[:c], mx["ngb_list"][k, :c])
            else:
                assert ov == 1
        small = np.where(mx["ngb_cnt"] < 256, mx["ngb_cnt"], -1)
        k = int(np.argmax(small))
        c = int(mx["ngb_cnt"][k])
        assert c > 2
        buf[:] = -7
        ov = L.g6_get_neighbour_list_(I(0), I(k), I(c), ctypes.byref(nn), P(buf, ctypes.c_int))
        assert ov == 1 and nn.value == c and np.all(buf[c:] == -7)
        buf[:] = -7
        ov = L.g6_get_neighbour_list_(I(0), I(k), I(c - 1), ctypes.byref(nn), P(buf, ctypes.c_int))
        assert ov == 1 and nn.value == c and buf[c - 2] != -7 and buf[c - 1] == -7
        # lasthalf (no NN) gives the same forces and leaves inn untouched
        L.g6calc_firsthalf_(I(0), I(n), I(ni), P(idi, ctypes.c_int), P(xi), P(vi), P(A), P(J), P(pot),
                            D(eps2), P(h2i))
        A2 = np.zeros((ni, 3)); J2 = np.zeros((ni, 3)); pot2 = np.zeros(ni)
        assert L.g6calc_lasthalf_(I(0), I(n), I(ni), P(idi, ctypes.c_int), P(xi), P(vi), D(eps2), P(h2i),
                                  P(A2), P(J2), P(pot2)) == 0
        assert np.array_equal(A, A2) and np.array_equal(J, J2) and np.array_equal(pot, pot2)
        # stored state: acc = 2 a2, jrk = 6 j6 (sapporoG6lib.cpp:103-104)
        L.get_j_part_data.argtypes = [ctypes.c_int, ctypes.c_int] + [c_dbl_p] * 6 + [c_dbl_p, c_dbl_p, c_int_p]
        out = [np.zeros(3) for _ in range(6)]
        mm = ctypes.c_double(); e2 = ctypes.c_double(); idv = ctypes.c_int()
        L.get_j_part_data(5, n, *[P(o) for o in out], ctypes.byref(mm), ctypes.byref(e2), ctypes.byref(idv))
        pos, vel, ac, jk, ppos, pvel = out
        assert np.array_equal(pos, x[5]) and np.array_equal(ac, 2 * (acc[5] / 2))
        assert np.array_equal(jk, 6 * (jrk[5] / 6)) and np.array_equal(ppos, px[5])
        assert mm.value == m[5] and idv.value == ids[5]
    finally:
        L.g6_close_(I(0))


def test_g6_overflow_and_error_paths(g6):
    _say("g6 overflow and error paths")
    L = g6
    n = 600
    rng = np.random.Generator(np.random.PCG64(32))
    x = rng.uniform(-1, 1, (n, 3)); v = np.zeros((n, 3)); m = np.full(n, 1.0 / n)
    ids = np.arange(n, dtype=np.int32)
    assert L.g6_open_special(0, None) == 0
    try:
        z = np.zeros((n, 3))
        _g6_load(L, m, x, v, ids, z, z, np.zeros(n))
        L.g6_set_ti_(I(0), D(0.0))
        ni = 4
        xi = np.ascontiguousarray(x[:ni]); vi = np.zeros((ni, 3)); idi = np.ascontiguousarray(ids[:ni])
        h2 = np.array([100.0, 0.01, 0.0, 1.0])        # i0: all 599 others are neighbours (>= 256)
        A = np.zeros((ni, 3)); J = np.zeros((ni, 3)); pot = np.zeros(ni); inn = np.zeros(ni, np.int32)
        L.g6calc_firsthalf_(I(0), I(n), I(ni), P(idi, ctypes.c_int), P(xi), P(vi), P(A), P(J), P(pot),
                            D(1e-4), P(h2))
        assert L.g6calc_lasthalf2_(I(0), I(n), I(ni), P(idi, ctypes.c_int), P(xi), P(vi), D(1e-4), P(h2),
                                   P(A), P(J), P(pot), P(inn, ctypes.c_int)) == 0
        assert L.g6_read_neighbour_list_(I(0)) == 1                  # overflow flagged
        buf = np.zeros(1024, np.int32); nn = ctypes.c_int(0)
        assert L.g6_get_neighbour_list_(I(0), I(0), I(1024), ctypes.byref(nn), P(buf, ctypes.c_int)) == 0
        assert nn.value == n - 1                                     # exact count past the 256 slots
        mx = oracle.neighbours_mixed(xi, idi, x, ids, h2)
        assert L.g6_get_neighbour_list_(I(0), I(1), I(1024), ctypes.byref(nn), P(buf, ctypes.c_int)) == 0
        assert nn.value == mx["ngb_cnt"][1] and np.array_equal(buf[:nn.value], mx["ngb_list"][1, :nn.value])
        assert L.g6_get_neighbour_list_(I(0), I(2), I(256), ctypes.byref(nn), P(buf, ctypes.c_int)) == 0
        assert nn.value == 0                                         # h2 = 0: no neighbours
        # ni = 0 and nj = 0 are silent no-ops (sapporohostclass.cpp:597, 718)
        L.g6calc_firsthalf_(I(0), I(n), I(0), P(idi, ctypes.c_int), P(xi), P(vi), P(A), P(J), P(pot),
                            D(1e-4), P(h2))
        assert L.g6calc_lasthalf_(I(0), I(0), I(ni), P(idi, ctypes.c_int), P(xi), P(vi), D(1e-4), P(h2),
                                  P(A), P(J), P(pot)) == 0
        # an address far beyond the grown capacity is rejected (sapporohostclass.cpp:456-466)
        bad = L.g6_set_j_particle_(I(0), I(10 ** 6), I(1), D(0), D(0), D(1.0), P(z[0]), P(z[0]), P(z[0]),
                                   P(z[0]), P(z[0]))
        assert bad != 0
    finally:
        L.g6_close_(I(0))
    # more i-particles than pipes: refused without touching the device
    with sa.Engine(sa.FOURTH, sa.DOUBLESINGLE) as eng:
        eng.set_j_particles(np.arange(4), np.arange(4), np.ones(4), np.eye(4, 3), np.zeros((4, 3)))
        big = 16385
        with pytest.raises(RuntimeError):
            eng.start_grav_calc(4, np.arange(big), np.zeros((big, 3)), np.zeros((big, 3)), 1e-4)


def test_sixth_order_abi():
    """initialize/set_j_particle/predict_all/pick_up_predictor_2/calc_force_on_predictors."""
    _say("6th-order abi")
    L = ctypes.CDLL(sa.LIB6_PATH)
    n = 3000
    m, x, v = ics.king(n, 6.0, seed=41)
    rng = np.random.Generator(np.random.PCG64(41))
    acc, jrk, snp, crk = (rng.normal(0, 0.3, (n, 3)) for _ in range(4))
    tj = -rng.uniform(0, 1.0 / 128, n)
    ids = (np.arange(n) + 7).astype(np.int32)
    L.initialize_special(1, (ctypes.c_int * 1)(0))
    L.set_j_particle.argtypes = [ctypes.c_int] + [c_dbl_p] * 6 + [ctypes.c_double, ctypes.c_double,
                                                                  ctypes.c_int, ctypes.c_double]
    for i in range(n):
        L.set_j_particle(i, *[P(np.ascontiguousarray(a[i])) for a in (x, v, acc, jrk, snp, crk)],
                         float(m[i]), float(tj[i]), int(ids[i]), 1e-6)
    ti = 1.0 / 256
    L.predict_all.argtypes = [ctypes.c_double, ctypes.c_int]
    L.predict_all(ti, n)
    px, pv, pa = oracle.predict(2, ti, tj, x, v, acc, jrk, snp, crk)
    L.pick_up_predictor_2.argtypes = [ctypes.c_int, c_int_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p]
    gp = np.zeros((n, 3)); gv = np.zeros((n, 3)); ga = np.zeros((n, 3))
    idv = ctypes.c_int(); mm = ctypes.c_double(); e2 = ctypes.c_double()
    for i in range(n):
        o = [np.zeros(3) for _ in range(3)]
        L.pick_up_predictor_2(i, ctypes.byref(idv), ctypes.byref(mm), ctypes.byref(e2), *[P(a) for a in o])
        gp[i], gv[i], ga[i] = o
        assert idv.value == ids[i] and mm.value == m[i] and e2.value == 1e-6
    assert np.array_equal(gp, px) and np.array_equal(gv, pv) and np.array_equal(ga, pa)
    eps2 = np.full(n, 1e-6)
    out = [np.zeros((n, 3)) for _ in range(4)]
    phi = np.zeros(n); nnb = np.zeros(n, np.int32); nnr2 = np.zeros(n)
    L.calc_force_on_predictors(n, n, P(ids, ctypes.c_int), P(gp), P(gv), P(ga), P(m), P(eps2),
                               *[P(o) for o in out], P(phi), P(nnb, ctypes.c_int), P(nnr2))
    accn, jrkn, snpn, crkn = out
    ref = oracle.force(2, gp, gv, ids, px, pv, m, ids, 1e-6, ai=ga, aj=pa, h2=np.zeros(n))
    for k, got in (("acc", accn), ("jrk", jrkn), ("snp", snpn), ("pot", phi)):
        assert np.median(oracle.rel_err(got, ref[k])) < 1e-12, k
    assert np.all(crkn == 0)
    nb, ds, found = oracle.api_nn(ref["nn_addr"], ref["nn_ds2"], ids)
    assert np.array_equal(nnb, nb) and np.array_equal(nnr2[found], ds[found])


def test_integrator_driver_energy():
    """The external GRAPE-6 Hermite-4 integrator, linked unchanged, conserves energy."""
    _say("shared-step integrator")
    exe = os.path.join(ROOT, "tests", "integrator", "hermite4_g6")
    assert os.path.exists(exe), "build tests/integrator first (__graft_entry__.build)"
    n = 1024
    m, x, v = ics.plummer(n, seed=51)
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), "h4_input_%d.bin" % os.getpid())
    with open(path, "wb") as f:
        np.array([n], np.int32).tofile(f)
        np.concatenate([m[:, None], x, v], axis=1).astype(np.float64).tofile(f)
    finals = {}
    for prec in ("mixed", "double"):
        env = dict(os.environ, SAPPORO2_AMD_PRECISION=prec)
        out = subprocess.run([exe, path, "0.25", "0.02", str(4.0 / n)], capture_output=True, text=True,
                             timeout=120, env=env)
        assert out.returncode == 0, out.stderr[-2000:]
        last = [l for l in out.stdout.splitlines() if l.startswith("final")][-1]
        finals[prec] = abs(float(last.split()[-1]))
        _say(f"integrator {prec}: {last}")
    os.remove(path)
    assert finals["double"] < 1e-5 and finals["mixed"] < 1e-5, finals


def test_wrapper_integrator_caller_side():
    """The integrator written against the header's caller side only (by-value wrappers,
    position-only multisend, board stubs; the call pattern of test_integrator.cpp:94-96,
    218, 238, 267): stubs return 0, a position-only load gives acc and pot bit-identical
    to the full load (jerk differs: vj = 0), and energy is conserved."""
    _say("caller-side wrapper integrator")
    exe = os.path.join(ROOT, "tests", "integrator", "hermite4_wrap_g6")
    assert os.path.exists(exe), "build tests/integrator first (__graft_entry__.build)"
    n = 1024
    m, x, v = ics.plummer(n, seed=53)
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), "h4w_input_%d.bin" % os.getpid())
    with open(path, "wb") as f:
        np.array([n], np.int32).tofile(f)
        np.concatenate([m[:, None], x, v], axis=1).astype(np.float64).tofile(f)
    for prec in ("mixed", "double"):
        env = dict(os.environ, SAPPORO2_AMD_PRECISION=prec)
        out = subprocess.run([exe, path, "0.125", "0.02", str(4.0 / n)], capture_output=True, text=True,
                             timeout=120, env=env)
        assert out.returncode == 0, out.stderr[-2000:]
        lines = out.stdout.splitlines()
        stubs = [l for l in lines if l.startswith("stubs")][0].split()
        fast = [l for l in lines if l.startswith("mxfast")][0].split()
        last = [l for l in lines if l.startswith("final")][-1]
        _say(f"wrapper integrator {prec}: {' '.join(stubs)} | {' '.join(fast)} | {last}")
        assert stubs[1] == "0" and stubs[3] == "16384", stubs
        assert fast[2] == "1" and fast[4] == "1", fast
        assert abs(float(last.split()[-1])) < 1e-5, last
    os.remove(path)


def test_block_timestep_integrator_energy():
    """The external block-timestep Hermite-4 integrator (individual power-of-two steps,
    active sets from single particles to all) conserves energy through the g6 ABI."""
    _say("block-timestep integrator")
    exe = os.path.join(ROOT, "tests", "integrator", "hermite4_block_g6")
    assert os.path.exists(exe), "build tests/integrator first (__graft_entry__.build)"
    n = 1024
    m, x, v = ics.plummer(n, seed=52)
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), "h4b_input_%d.bin" % os.getpid())
    with open(path, "wb") as f:
        np.array([n], np.int32).tofile(f)
        np.concatenate([m[:, None], x, v], axis=1).astype(np.float64).tofile(f)
    finals = {}
    for prec in ("mixed", "double"):
        env = dict(os.environ, SAPPORO2_AMD_PRECISION=prec)
        out = subprocess.run([exe, path, "0.25", "0.02", str(4.0 / n)], capture_output=True, text=True,
                             timeout=120, env=env)
        assert out.returncode == 0, out.stderr[-2000:]
        blk = [l for l in out.stdout.splitlines() if l.startswith("blocks")][-1].split()
        last = [l for l in out.stdout.splitlines() if l.startswith("final")][-1]
        finals[prec] = abs(float(last.split()[-1]))
        _say(f"block integrator {prec}: {' '.join(blk)} {last}")
        nblocks, mean_active, min_active = int(blk[1]), float(blk[3]), int(blk[5])
        assert nblocks > 16 and mean_active < n / 2 and min_active < 16, blk   # real block steps
    os.remove(path)
    assert finals["double"] < 1e-5 and finals["mixed"] < 1e-5, finals


def test_hermite6_block_integrator_energy():
    """The external block-timestep Hermite-6 integrator (predict_all, pick_up_predictor_2,
    calc_force_on_predictors, crackle from the quintic through acc, jerk and snap at both
    ends of a step) conserves energy through the 6th-order ABI.  The same driver linked
    against the oracle's FP64 sums gives dE/E = 9.38e-10 on this input (162 blocks)."""
    _say("hermite-6 block-timestep integrator")
    exe = os.path.join(ROOT, "tests", "integrator", "hermite6_block_6th")
    assert os.path.exists(exe), "build tests/integrator first (__graft_entry__.build)"
    n = 1024
    m, x, v = ics.plummer(n, seed=52)
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), "h6b_input_%d.bin" % os.getpid())
    with open(path, "wb") as f:
        np.array([n], np.int32).tofile(f)
        np.concatenate([m[:, None], x, v], axis=1).astype(np.float64).tofile(f)
    finals = {}
    for prec in ("double", "mixed"):
        env = dict(os.environ, SAPPORO2_AMD_PRECISION=prec)
        out = subprocess.run([exe, path, "0.25", "0.1", str(4.0 / n)], capture_output=True, text=True,
                             timeout=100, env=env)
        assert out.returncode == 0, out.stderr[-2000:]
        blk = [l for l in out.stdout.splitlines() if l.startswith("blocks")][-1].split()
        last = [l for l in out.stdout.splitlines() if l.startswith("final")][-1]
        finals[prec] = abs(float(last.split()[-1]))
        _say(f"hermite-6 block integrator {prec}: {' '.join(blk)} {last}")
        nblocks, mean_active, min_active = int(blk[1]), float(blk[3]), int(blk[5])
        assert nblocks > 16 and mean_active < n / 2 and min_active < 16, blk
    os.remove(path)
    assert finals["double"] < 1e-8 and finals["mixed"] < 1e-7, finals


@pytest.mark.parametrize("prec", [sa.DOUBLESINGLE, sa.DOUBLE])
def test_full_size_properties(prec):
    """N = 131072: bitwise reproducible, Newton's third law, exact NN/lists on a sample."""
    _say(f"full size prec={prec}")
    n = 131072
    m, x, v, ids, h2 = plummer_system(n, seed=61)
    eps2 = 1e-6
    with sa.Engine(sa.FOURTH, prec) as eng:
        eng.set_j_particles(np.arange(n), ids, m, x, v)
        eng.set_time(0.0)
        r1 = eng.calc(n, ids, x, v, eps2, h2)
        r2 = eng.calc(n, ids, x, v, eps2, h2)
        for k in ("acc", "jrk", "pot", "nnb", "dsmin"):
            assert np.array_equal(r1[k], r2[k], equal_nan=True), k
        tot = (m[:, None] * r1["acc"]).sum(0)
        scale = (m[:, None] * np.abs(r1["acc"])).sum(0)
        assert np.all(np.abs(tot) < (1e-6 if prec == sa.DOUBLESINGLE else 1e-12) * scale)
        rng = np.random.Generator(np.random.PCG64(62))
        sel = np.sort(rng.choice(n, 256, replace=False))
        eng.start_grav_calc(n, ids[sel], x[sel], v[sel], eps2, h2[sel])
        rs = eng.get_grav_results(n, len(sel))
        from _util import check_forces, check_nn_lists
        ref = oracle.force(1, x[sel], v[sel], ids[sel], x, v, m, ids, eps2, h2=h2[sel])
        check_forces(rs, ref, 1, 1e-6 if prec == sa.DOUBLESINGLE else 1e-12, label="full")
        check_nn_lists(eng, rs, prec == sa.DOUBLE, x[sel], ids[sel], x, ids, h2[sel], label="full")
