"""GRAPE-5 C-ABI (include/g5lib.h, reference sapporoG5lib.cpp:40-218) against the FP64
oracle: second-order force in single precision on float-rounded positions, no
self-exclusion (the own mass adds -m/eps to the potential), calls larger than the
16384 pipelines split in chunks, the asynchronous set_xi/run/get_force path, and the
unset-softening error path."""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

import oracle
import sapporo2_amd as sa
from sapporo2_amd import ics

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def P(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.fixture(scope="module")
def g5():
    L = ctypes.CDLL(sa.LIB_PATH)
    L.g5_set_eps_to_all.argtypes = [ctypes.c_double]
    L.g5_set_eps2_to_all.argtypes = [ctypes.c_double]
    return L


def _reference(x, m, xi, eps2):
    """FP64 oracle on the float-rounded positions; ids -1 on the i side: no pair excluded."""
    xj = x.astype(np.float32).astype(np.float64)
    xf = xi.astype(np.float32).astype(np.float64)
    z = np.zeros_like(xf)
    return oracle.force(1, xf, z, np.full(len(xf), -1, np.int32), xj, np.zeros_like(xj), m,
                        np.arange(len(xj), dtype=np.int32), eps2)


def test_g5_force_and_chunks(g5):
    L = g5
    n = 20000
    m, x, _ = ics.plummer(n, seed=71)
    x = np.ascontiguousarray(x)
    _say("g5 open")
    assert L.g5_open_special(0, None) == 0
    try:
        assert L.g5_get_number_of_pipelines() == 16384
        L.g5_set_range(ctypes.c_double(-1.0), ctypes.c_double(1.0), ctypes.c_double(0.0))
        L.g5_set_xmj(0, n, P(x), P(m))
        L.g5_set_n(n)
        eps = 0.01
        L.g5_set_eps_to_all(eps)
        # every particle as an i-particle: 20000 > 16384, two chunks
        a = np.zeros((n, 3)); phi = np.zeros(n)
        _say("g5 calculate_force_on_x, ni=20000")
        L.g5_calculate_force_on_x(P(x), P(a), P(phi), n)
        sub = np.arange(0, n, 10)
        ref = _reference(x, m, x[sub], eps * eps)
        ea = oracle.rel_err(a[sub], ref["acc"])
        ep = oracle.rel_err(phi[sub], ref["pot"])
        _say(f"g5 acc median {np.median(ea):.2e} pot median {np.median(ep):.2e}")
        assert np.median(ea) < 1e-5 and np.median(ep) < 1e-5
        # self term: a particle alone with itself has pot = -m/eps
        assert np.all(phi < 0)
        # asynchronous path on a subset, softening given as eps^2
        L.g5_set_eps2_to_all(1e-4)
        xi = np.ascontiguousarray(x[:1000] + 1e-3)
        L.g5_set_xi(1000, P(xi))
        L.g5_run()
        a2 = np.zeros((1000, 3)); phi2 = np.zeros(1000)
        L.g5_get_force(1000, P(a2), P(phi2))
        ref2 = _reference(x, m, xi, 1e-4)
        assert np.median(oracle.rel_err(a2, ref2["acc"])) < 1e-5
        assert np.median(oracle.rel_err(phi2, ref2["pot"])) < 1e-5
        # nj restricts the j-particles to [0, nj)
        L.g5_set_n(5000)
        a3 = np.zeros((100, 3)); phi3 = np.zeros(100)
        L.g5_calculate_force_on_x(P(np.ascontiguousarray(x[:100])), P(a3), P(phi3), 100)
        ref3 = _reference(x[:5000], m[:5000], x[:100], 1e-4)
        assert np.median(oracle.rel_err(a3, ref3["acc"])) < 1e-5
        assert L.g5_get_jmemsize() > 0 and L.g5_get_number_of_real_pipelines() == 1
    finally:
        L.g5_close()


def test_g5_requires_softening(g5):
    L = g5
    assert L.g5_open_special(0, None) == 0
    try:
        n = 64
        m, x, _ = ics.plummer(n, seed=72)
        x = np.ascontiguousarray(x)
        L.g5_set_xmj(0, n, P(x), P(m))
        L.g5_set_n(n)
        a = np.full((n, 3), 7.0); phi = np.full(n, 7.0)
        L.g5_calculate_force_on_x(P(x), P(a), P(phi), n)   # eps never set: refused, outputs untouched
        assert np.all(a == 7.0) and np.all(phi == 7.0)
    finally:
        L.g5_close()
