"""CPU oracle for the direct-summation hot path.

TEST INFRASTRUCTURE ONLY: the reference results for ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``.  The product library never imports it.

Thin ctypes bindings over ``oracle/liboracle.so`` (built from ``oracle.c`` by the
Makefile beside it) plus the host-side API semantics of the reference's g6 result
readback (``lib/src/sapporohostclass.cpp:705-880``):

* ``predict``            -- FP64 predictor (``kernels.cu:332-383``), bit-exact contract.
* ``force``              -- FP64 acc/pot/jerk[/snap] + NN + lists (``sapporohostclass.cpp:1373-1438``,
                            ``kernels.cu:565-667``, ``sapporoYeblib.cpp:57-86``).
* ``neighbours_mixed``   -- NN + lists under the FP32 (mixed-precision) distance contract.
* ``api_nn``             -- the readback rule: nnb/dsmin reported only when dsmin < 1e10
                            (``sapporohostclass.cpp:730-775``).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ORDER_G5, ORDER_H4, ORDER_H6 = 0, 1, 2
NGB_CAP = 256          # reference NGB_PB (lib/include/defines.h:9)
NN_THRESHOLD = 1.0e10  # reference host threshold (sapporohostclass.cpp:732, 767)

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


def build_native(outdir):
    """Compile oracle.c for this host's CPU (``-O3 -march=native``, no FP contraction)
    into ``outdir`` and load it in place of the portable build.  Same arithmetic: the
    restatement has no reductions the compiler may reorder without -ffast-math, and
    fma/sqrt are correctly rounded either way.  Returns the compile command, or None
    (the portable x86-64-v3 build stays loaded) when the compiler is unavailable."""
    import subprocess
    global _LIB
    out = os.path.join(outdir, "liboracle_native.so")
    cmd = ["gcc", "-O3", "-march=native", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
           "-shared", "-o", out, os.path.join(_HERE, "oracle.c"), "-lm"]
    try:
        subprocess.check_call(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    except (OSError, subprocess.CalledProcessError):
        return None
    _LIB = None
    lib(out)
    return " ".join(cmd[:6] + ["oracle.c"])


def lib(path=None):
    """Load liboracle.so (or the library at ``path``), building it with make if it is missing."""
    global _LIB
    if _LIB is None:
        if path is None:
            path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", _HERE])
        L = ctypes.CDLL(path)
        L.orc_predict.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp,
                                  _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_predict.restype = None
        L.orc_force.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _ip,
                                ctypes.c_int, _dp, _dp, _dp, _dp, _ip, ctypes.c_double,
                                _dp, _dp, _dp, _dp, _ip, _dp, _ip, _ip, ctypes.c_int, ctypes.c_int]
        L.orc_force.restype = None
        L.orc_neighbours_mixed.argtypes = [ctypes.c_int, _dp, _dp, _ip, ctypes.c_int, _dp, _ip,
                                           _ip, _fp, _ip, _ip, ctypes.c_int, ctypes.c_int]
        L.orc_neighbours_mixed.restype = None
        L.orc_force_plain.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _ip, ctypes.c_int,
                                      _dp, _dp, _dp, _dp, _ip, ctypes.c_double, _dp, _dp, _dp, _dp,
                                      ctypes.c_int]
        L.orc_force_plain.restype = None
        L.orc_max_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def max_threads() -> int:
    return int(lib().orc_max_threads())


def predict(order, ti, tj, x, v, a, j, s=None, c=None):
    """FP64 predictor; returns (px, pv, pa) with pa None below SIXTH."""
    tj, x, v, a, j, s, c = map(_d, (tj, x, v, a, j, s, c))
    n = x.shape[0]
    px = np.empty((n, 3)); pv = np.empty((n, 3))
    pa = np.empty((n, 3)) if order >= ORDER_H6 else None
    lib().orc_predict(n, order, float(ti), _p(tj, _dp), _p(x, _dp), _p(v, _dp), _p(a, _dp),
                      _p(j, _dp), _p(s, _dp), _p(c, _dp), _p(px, _dp), _p(pv, _dp), _p(pa, _dp))
    return px, pv, pa


def force(order, xi, vi, idi, xj, vj, mj, idj, eps2, ai=None, aj=None, h2=None,
          cap=NGB_CAP, nthreads=0, lists=True):
    """FP64 evaluation.  Returns a dict with acc, jrk, snp, pot, nn_addr, nn_ds2,
    ngb_cnt and ngb_list (ni x cap, sorted, padded with -1)."""
    xi, vi, ai, xj, vj, aj, mj, h2 = map(_d, (xi, vi, ai, xj, vj, aj, mj, h2))
    idi, idj = _i(idi), _i(idj)
    ni, nj = xi.shape[0], xj.shape[0]
    acc = np.zeros((ni, 3)); jrk = np.zeros((ni, 3)); pot = np.zeros(ni)
    snp = np.zeros((ni, 3)) if order >= ORDER_H6 else None
    nn_addr = np.empty(ni, np.int32); nn_ds2 = np.empty(ni)
    cnt = np.zeros(ni, np.int32)
    lst = np.full((ni, cap), -1, np.int32) if (lists and h2 is not None) else None
    lib().orc_force(order, ni, _p(xi, _dp), _p(vi, _dp), _p(ai, _dp), _p(h2, _dp), _p(idi, _ip),
                    nj, _p(xj, _dp), _p(vj, _dp), _p(aj, _dp), _p(mj, _dp), _p(idj, _ip),
                    float(eps2), _p(acc, _dp), _p(jrk, _dp), _p(snp, _dp), _p(pot, _dp),
                    _p(nn_addr, _ip), _p(nn_ds2, _dp), _p(cnt, _ip), _p(lst, _ip),
                    cap if lst is not None else 0, int(nthreads))
    return dict(acc=acc, jrk=jrk, snp=snp, pot=pot, nn_addr=nn_addr, nn_ds2=nn_ds2,
                ngb_cnt=cnt, ngb_list=lst)


def force_plain(order, xi, vi, idi, xj, vj, mj, idj, eps2, ai=None, aj=None, nthreads=0):
    """FP64 forces only (the timed CPU baseline)."""
    xi, vi, ai, xj, vj, aj, mj = map(_d, (xi, vi, ai, xj, vj, aj, mj))
    idi, idj = _i(idi), _i(idj)
    ni, nj = xi.shape[0], xj.shape[0]
    acc = np.zeros((ni, 3)); jrk = np.zeros((ni, 3)); pot = np.zeros(ni)
    snp = np.zeros((ni, 3)) if order >= ORDER_H6 else None
    lib().orc_force_plain(order, ni, _p(xi, _dp), _p(vi, _dp), _p(ai, _dp), _p(idi, _ip), nj,
                          _p(xj, _dp), _p(vj, _dp), _p(aj, _dp), _p(mj, _dp), _p(idj, _ip),
                          float(eps2), _p(acc, _dp), _p(jrk, _dp), _p(snp, _dp), _p(pot, _dp),
                          int(nthreads))
    return dict(acc=acc, jrk=jrk, snp=snp, pot=pot)


def neighbours_mixed(xi, idi, xj, idj, h2, cap=NGB_CAP, nthreads=0, lists=True):
    """NN + lists under the FP32 distance contract (mixed-precision kernel)."""
    xi, xj, h2 = map(_d, (xi, xj, h2))
    idi, idj = _i(idi), _i(idj)
    ni, nj = xi.shape[0], xj.shape[0]
    nn_addr = np.empty(ni, np.int32); nn_ds2 = np.empty(ni, np.float32)
    cnt = np.zeros(ni, np.int32)
    lst = np.full((ni, cap), -1, np.int32) if (lists and h2 is not None) else None
    lib().orc_neighbours_mixed(ni, _p(xi, _dp), _p(h2, _dp), _p(idi, _ip), nj, _p(xj, _dp),
                               _p(idj, _ip), _p(nn_addr, _ip), _p(nn_ds2, _fp), _p(cnt, _ip),
                               _p(lst, _ip), cap if lst is not None else 0, int(nthreads))
    return dict(nn_addr=nn_addr, nn_ds2=nn_ds2, ngb_cnt=cnt, ngb_list=lst)


def api_nn(nn_addr, nn_ds2, idj):
    """Apply the reference readback rule to raw NN results: returns (nnb, dsmin, found)
    where found[i] is False when the minimum is not below 1e10 (then nnb = 0 and the
    caller's dsmin is left untouched; sapporohostclass.cpp:730-775).  dsmin is the
    float-rounded distance, as the device returns it (kernels.cu:999-1002)."""
    nn_addr = np.asarray(nn_addr)
    ds = np.asarray(nn_ds2).astype(np.float32).astype(np.float64)
    found = (nn_addr >= 0) & (ds < NN_THRESHOLD)
    idj = np.asarray(idj)
    nnb = np.where(found, idj[np.clip(nn_addr, 0, None)], 0).astype(np.int32)
    return nnb, np.where(found, ds, np.nan), found


def rel_err(x, ref):
    """Per-particle relative error |x - ref| / |ref| over the last axis
    (test_gravity_block.cpp:359-362)."""
    x = np.asarray(x, np.float64); ref = np.asarray(ref, np.float64)
    if x.ndim == 1:
        return np.abs(x - ref) / np.abs(ref)
    return np.sqrt(np.sum((x - ref) ** 2, axis=-1) / np.sum(ref ** 2, axis=-1))
