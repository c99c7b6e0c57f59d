"""The reference's Signal=0 initial state generated by the library
(msc_ini_state, host code, no GPU): the lattice against v25/v251/run1's
iniPos.dat, and the whole initialisation (lattice, 1000 GCAmix moves,
ini_mom) against the oracle's, whose start reproduces the run's output.dat
(tests/test_oracle_fixtures.py)."""
import ctypes
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
L251 = math.pow(500.0 / ((1.0 * 1.0) * 0.0013), 1.0 / 3)


@pytest.fixture(scope="module")
def msc():
    import mdspherocyl_amd as m
    m.load()
    return m


def test_lattice_matches_ini_pos_dat(msc):
    st = msc.ini_state(500, (L251,) * 3, 9.0, 1.0, 1.0, 1.0, nmix=0)
    ref = open(os.path.join(GOLD, "v251_run1_iniPos.txt")).read().splitlines()
    for i in range(500):
        assert ["%g" % v for v in st["r"][i]] == ref[i].split()[:3], "particle %d" % i
    assert np.array_equal(st["ori"], np.tile(np.eye(3), (500, 1, 1)))


def test_initial_state_equals_oracle_bitwise(msc):
    """r, ori, p, l and the advanced generator equal the oracle's
    o_run_init(seed_init) state of v25/v251/run1 bit for bit."""
    st = msc.ini_state(500, (L251,) * 3, 9.0, 1.0, 1.0, 1.0)
    S, P = H.v251_state()
    R = O.ORun(S, P, seed_init=True)
    try:
        for k in ("r", "p", "l", "ori"):
            assert np.array_equal(st[k], getattr(S, k)), k
        mt = np.ctypeslib.as_array(R.R.ran.mt, shape=(624,))
        assert st["mt_index"] == R.R.ran.mti and np.array_equal(st["mt"], mt)
    finally:
        R.free()
    # momenta sum to zero as ini_mom leaves them: the last particle cancels the rest
    assert np.all(np.abs(st["p"].sum(0)) <= 1e-12) and np.all(np.abs(st["l"].sum(0)) <= 1e-12)


def test_ini_state_odd_count_and_errors(msc):
    """An odd particle count takes the reference's extra draw for the last
    particle (compared with the oracle); a box too small for the lattice and
    a second species are refused."""
    N, L = 101, (12.5, 12.5, 60.0)
    st = msc.ini_state(N, L, 2.0, 1.0, 1.5, 1.3, nmix=50)
    S = O.OSys(N, 1, L=L)
    S.len[:], S.dia[:], S.mass[:] = 2.0, 1.0, 1.5
    S.set_inertia()
    g = O.COmt()
    O.lib().omt_seed(ctypes.byref(g), 1)
    assert O.lib().o_ini_pos(S.c()) == 0
    for _ in range(50):
        O.lib().o_gcamix(S.c(), ctypes.byref(g))
    assert O.lib().o_ini_mom(S.c(), 1.3, ctypes.byref(g)) == 0
    for k in ("r", "p", "l", "ori"):
        assert np.array_equal(st[k], getattr(S, k)), k
    with pytest.raises(msc.Overlap):
        msc.ini_state(400, (4.0, 4.0, 4.0), 2.0, 1.0, 1.0, 1.0, nmix=0)
    lib = msc.load()
    mt, idx = msc.mt19937_state(1)
    arr = [np.zeros(6) for _ in range(3)]
    out = [np.zeros(18), np.zeros(18), np.zeros(18), np.zeros(54)]
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    counts = np.array([3, 3], dtype=np.int32)
    rc = lib.msc_ini_state(6, 2, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), dp(np.full(3, 10.0)),
                           *[dp(a + 1) for a in arr], 1.0, 0, mt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                           ctypes.byref(ctypes.c_int(idx)), *[dp(a) for a in out])
    assert rc == msc.MSC_E_ARG if hasattr(msc, "MSC_E_ARG") else rc != 0
