"""The allocation-free entry point sapamd_grav_calc_noalloc against the regular one
(sapamd_start_grav_calc + sapamd_get_grav_results), through the C-ABI.

Two calls in a row, the second with a smaller active set and a prediction pending inside
the allocation-free call: results bit-identical to the regular entry point's (same launch
shape, same kernels) and within tolerance of the FP64 oracle.  A call too large for the
scratch array at the regular launch shape coarsens its j-split: within tolerance of the
regular results, nearest neighbours identical.  Error paths: too many i-particles, staged
updates pending.  Device buffers for the inputs and outputs are the test's own."""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

import oracle
import sapporo2_amd as sa
from sapporo2_amd import ics
from _util import check_forces

pytestmark = pytest.mark.gpu
TOL = {sa.DOUBLESINGLE: 1e-6, sa.DOUBLE: 1e-12}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Hip:
    """Just enough of the HIP runtime for the test's own device buffers."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.bufs = []

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(nbytes, 1))) == 0
        self.bufs.append(p)
        return p

    def h2d(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        assert self.L.hipMemcpy(dst, arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(arr.nbytes), 1) == 0

    def d2h(self, arr, src):
        assert self.L.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), src, ctypes.c_size_t(arr.nbytes), 2) == 0

    def free_all(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _noalloc(eng, hip, nj, ids, xi, vi, eps2, stream=None):
    """Runs the allocation-free call on device copies of the inputs; returns the result
    dict in the regular entry point's form (nnb / dsmin with its 1e10 read-back rule)."""
    L = eng._L
    L.sapamd_grav_calc_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
    ni = len(ids)
    nf = 13 if eng.order >= sa.SIXTH else 10
    d_id = hip.malloc(4 * ni)
    d_in = hip.malloc(8 * 6 * ni)
    d_out = hip.malloc(8 * nf * ni)
    hip.h2d(d_id, np.asarray(ids, np.int32))
    hip.h2d(d_in, np.concatenate([np.ravel(xi), np.ravel(vi)]).astype(np.float64))
    rc = L.sapamd_grav_calc_noalloc(eng._h, nj, ni, d_id, d_in, eps2, d_out, stream)
    if rc != 0:
        return rc
    eng.synchronize() if stream is None else hip.L.hipStreamSynchronize(stream)
    out = np.empty((nf, ni))
    hip.d2h(out, d_out)
    ds, nn = out[7], out[8]
    found = ds < 1.0e10
    return dict(acc=out[0:3].T.copy(), pot=out[3].copy(), jrk=out[4:7].T.copy(),
                nnb=np.where(found, nn, 0).astype(np.int32), dsmin=np.where(found, ds, np.nan))


@pytest.mark.parametrize("prec", [sa.DOUBLESINGLE, sa.DOUBLE])
def test_noalloc_matches_regular_entry(prec):
    n, eps2 = 65536, 1e-6
    m, x, v = ics.plummer(n, seed=301)
    ids = np.arange(n, dtype=np.int32)
    rng = np.random.Generator(np.random.PCG64(302))
    acc, jrk = rng.normal(0, 0.3, (n, 3)), rng.normal(0, 0.3, (n, 3))
    tj = -rng.uniform(0, 1.0 / 256, n)
    hip = Hip()
    L = sa.lib()
    L.sapamd_noalloc_scratch_bytes.restype = ctypes.c_longlong
    _say(f"noalloc prec={prec}: scratch {L.sapamd_noalloc_scratch_bytes() >> 20} MiB")
    try:
        with sa.Engine(sa.FOURTH, prec) as eng:
            eng.set_j_particles(np.arange(n), ids, m, x, v, acc, jrk, tj=tj)
            # call 1: the regular entry point flushes and predicts; then the same call allocation-free
            t1 = 1.0 / 512
            eng.set_time(t1)
            px, pv, _ = oracle.predict(1, t1, tj, x, v, acc, jrk)
            s1 = np.sort(rng.choice(n, 3000, replace=False))
            eng.start_grav_calc(n, ids[s1], px[s1], pv[s1], eps2, None)
            reg = eng.get_grav_results(n, len(s1))
            got = _noalloc(eng, hip, n, ids[s1], px[s1], pv[s1], eps2)
            for k in ("acc", "jrk", "pot", "nnb", "dsmin"):
                assert np.array_equal(got[k], reg[k], equal_nan=True), f"call 1 {k}"
            ref = oracle.force(1, px[s1], pv[s1], ids[s1], px, pv, m, ids, eps2)
            check_forces(got, ref, 1, TOL[prec], label=f"noalloc call 1 prec={prec}")
            # call 2, smaller, with the prediction to t2 pending inside the allocation-free call,
            # on a stream of the caller's
            t2 = 1.0 / 256
            eng.set_time(t2)
            px, pv, _ = oracle.predict(1, t2, tj, x, v, acc, jrk)
            s2 = np.sort(rng.choice(n, 700, replace=False))
            stream = ctypes.c_void_p()
            assert hip.L.hipStreamCreate(ctypes.byref(stream)) == 0
            got = _noalloc(eng, hip, n, ids[s2], px[s2], pv[s2], eps2, stream)
            hip.L.hipStreamDestroy(stream)
            eng.start_grav_calc(n, ids[s2], px[s2], pv[s2], eps2, None)
            reg = eng.get_grav_results(n, len(s2))
            for k in ("acc", "jrk", "pot", "nnb", "dsmin"):
                assert np.array_equal(got[k], reg[k], equal_nan=True), f"call 2 {k}"
            ref = oracle.force(1, px[s2], pv[s2], ids[s2], px, pv, m, ids, eps2)
            check_forces(got, ref, 1, TOL[prec], label=f"noalloc call 2 prec={prec}")
            nnb, dsmin, found = oracle.api_nn(*(oracle.neighbours_mixed(px[s2], ids[s2], px, ids, np.zeros(len(s2)),
                                                                        lists=False)[k]
                                                for k in ("nn_addr", "nn_ds2")), ids) if prec == sa.DOUBLESINGLE \
                else oracle.api_nn(ref["nn_addr"], ref["nn_ds2"], ids)
            assert np.array_equal(got["nnb"], nnb) and np.array_equal(got["dsmin"][found], dsmin[found])
            # a full 16384-particle call: its regular j-split does not fit, the coarser one agrees
            s3 = np.sort(rng.choice(n, 16384, replace=False))
            eng.start_grav_calc(n, ids[s3], px[s3], pv[s3], eps2, None)
            reg = eng.get_grav_results(n, len(s3))
            got = _noalloc(eng, hip, n, ids[s3], px[s3], pv[s3], eps2)
            check_forces(got, reg, 1, TOL[prec], label=f"noalloc 16384 vs regular prec={prec}")
            assert np.array_equal(got["nnb"], reg["nnb"])
            assert np.array_equal(got["dsmin"], reg["dsmin"], equal_nan=True)
            # error paths: too many i-particles; staged updates pending
            assert _noalloc(eng, hip, n, np.arange(16385), np.zeros((16385, 3)), np.zeros((16385, 3)), eps2) == -1
            eng.set_j_particles([5], ids[[5]], m[[5]], x[[5]], v[[5]])
            assert _noalloc(eng, hip, n, ids[s2], px[s2], pv[s2], eps2) == -1
            eng.force_prediction(n)
            assert isinstance(_noalloc(eng, hip, n, ids[s2], px[s2], pv[s2], eps2), dict)
    finally:
        hip.free_all()
