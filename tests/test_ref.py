"""The oracle and the library against the reference's own code where it
builds here (oracle/_ref/libref.so from oracle/build_ref.py: xorshift64,
the Stejskal-Tanner least-squares fit, accumulated linspace / logspace).
Skipped when that library is absent (it needs the reference tree to build).
"""
import ctypes as ct
import math
import os

import numpy as np
import pytest

import oracle as O
import randomwalk_nmr_amd as R
from oracle import laplace as OL
from randomwalk_nmr_amd import laplace as RL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref/libref.so not built")


@pytest.fixture(scope="module")
def ref():
    L = ct.CDLL(REF_LIB)
    P, D, I = ct.c_void_p, ct.c_double, ct.c_int
    L.ref_xorshift64_sequence.argtypes = [ct.c_uint64, I, P]
    L.ref_linspace.argtypes = [D, D, I, I, P]
    L.ref_logspace.argtypes = [D, D, I, I, P]
    L.ref_stejskal_tanner.argtypes = [P, P, I, I, ct.POINTER(D), ct.POINTER(D), ct.POINTER(D)]
    return L


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


def test_xorshift64_matches_reference(ref):
    rng = np.random.default_rng(0)
    seeds = [1, 2, 0x8000000000000000, 0xFFFFFFFFFFFFFFFF] + [int(v) for v in rng.integers(1, 2**63, 200)]
    for s in seeds:
        want = np.zeros(64, np.uint64)
        ref.ref_xorshift64_sequence(s, 64, _p(want))
        got, cur = [], s
        for _ in range(64):
            cur = O.xorshift64(cur)
            got.append(cur)
        assert np.array_equal(np.array(got, np.uint64), want), hex(s)


@pytest.mark.parametrize("start,end,size,final", [(-4.0, 2.0, 512, True), (math.log10(0.1), math.log10(1e4), 128, True),
                                                  (0.0, 1.0, 7, False), (2.5, -1.25, 33, True)])
def test_accumulated_spaces_match_reference(ref, start, end, size, final):
    lin = np.zeros(size)
    log = np.zeros(size)
    ref.ref_linspace(start, end, size, int(final), _p(lin))
    ref.ref_logspace(start, end, size, int(final), _p(log))
    assert np.array_equal(OL.linspace_accumulated(start, end, size, final), lin)
    assert np.array_equal(OL.logspace_accumulated(start, end, size, final), log)


def test_library_t2_bins_and_lambdas_match_reference(ref):
    times = np.arange(1, 2001) * 0.2
    inv = RL.Inverter(times, num_lambdas=512)
    bins = np.zeros(128)
    ref.ref_logspace(math.log10(0.1), math.log10(1e4), 128, 1, _p(bins))
    assert np.array_equal(inv.t2_bins, bins)
    inv.find_best_lambda(0.4 * np.exp(-times / 30.0) + 0.6 * np.exp(-times / 500.0))
    lams = np.zeros(512)
    ref.ref_logspace(-4.0, 2.0, 512, 1, _p(lams))
    assert np.array_equal(inv.lcurve()["lambdas"], lams)


@pytest.mark.parametrize("n", [2, 5, 64])
def test_stejskal_tanner_matches_reference(ref, n):
    """Increasing |k| (the reference's gradient lists), no intercept
    (NMR_pfgse.cpp:680): the library and the oracle give the reference's D
    and standard error."""
    L = R.load()
    rng = np.random.default_rng(n)
    for trial in range(20):
        k = np.sort(rng.uniform(0.0, 3.0, n))
        k[0] = 0.0
        rhs = -(k ** 2) * rng.uniform(0.5, 40.0)
        lhs = rhs * rng.uniform(1.0, 2.5) + 1e-3 * rng.standard_normal(n)
        D, E, A = ct.c_double(), ct.c_double(), ct.c_double()
        ref.ref_stejskal_tanner(_p(rhs), _p(lhs), n, 0, ct.byref(D), ct.byref(E), ct.byref(A))
        assert A.value == 0.0
        b, e = OL.stejskal_tanner(rhs, lhs)
        assert abs(b - D.value) <= 1e-15 * abs(D.value) and abs(e - E.value) <= 1e-12 * max(E.value, 1e-300)
        d2, e2 = ct.c_double(), ct.c_double()
        assert L.rwnmr_stejskal_tanner(_p(rhs), _p(lhs), n, ct.byref(d2), ct.byref(e2)) == 0
        assert abs(d2.value - D.value) <= 1e-15 * abs(D.value)
        assert abs(e2.value - E.value) <= 1e-12 * max(E.value, 1e-300)
