"""The reference-side adapter on the GPU: the force routines of one reference
step called through mdspherocyl_amd/adapter/msc_adapter.hpp on the test
stand-ins of the reference's containers (tests/adapter), against the oracle.

This covers what the adapter adds to the C ABI: gathering part[species][i]
species-major, recovering the body-frame dipole constraints from the lab-frame
dipoles and get_lev, adding F and T back with add_F / add_T, filling the
caller's nblist, uploading a list of the caller's own (msc_set_neigh_list),
and creating the context with alpha and K from tabul / wave_v.

Tolerances: the neighbour list's pair set is exact; forces and torques of the
whole step within 1e-10 of the oracle per particle relative to the rms
magnitude (reciprocal sums by separable phase factors); energies 1e-10
relative; the step over an uploaded list equals the step over the list
neigh_list filled to 1e-12.
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

PROBE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "adapter", "libadapter_probe.so")
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def probe():
    import mdspherocyl_amd as m
    m.load()
    if not os.path.exists(PROBE):
        pytest.fail("tests/adapter/libadapter_probe.so is not built (__graft_entry__.build())")
    L = ctypes.CDLL(PROBE)
    L.probe_create.restype = ctypes.c_void_p
    L.probe_create.argtypes = [ctypes.c_int, _ip, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int] + [_dp] * 10
    L.probe_destroy.argtypes = [ctypes.c_void_p]
    L.probe_neigh_list.restype = ctypes.c_long
    L.probe_neigh_list.argtypes = [ctypes.c_void_p, ctypes.c_double]
    L.probe_give_list.argtypes = [ctypes.c_void_p, _ip, _ip]
    L.probe_step.restype = ctypes.c_int
    L.probe_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                             _dp, _dp, _dp, _dp]
    L.probe_list.argtypes = [ctypes.c_void_p, _ip, _ip]
    return L


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def _create(L, S, alpha, K, B):
    keep = []

    def d(x):
        a, p = _d(x)
        keep.append(a)
        return p
    counts, cp = _i(S.counts)
    keep.append(counts)
    h = L.probe_create(len(S.counts), cp, S.nd, d(S.L), d(B), alpha, K, d(S.r), d(S.p), d(S.l), d(S.ori),
                       d(S.len), d(S.dia), d(S.mass), d(S.dpc), d(S.doc), d(S.dmu))
    return h, keep


def _step(L, h, N, use_given, r_cut, K, alpha):
    E, PT = np.zeros(6), np.zeros(27)
    F, T = np.zeros((N, 3)), np.zeros((N, 3))
    rc = L.probe_step(h, use_given, r_cut, K, alpha, E.ctypes.data_as(_dp),
                      PT.ctypes.data_as(_dp), F.ctypes.data_as(_dp), T.ctypes.data_as(_dp))
    return rc, E, PT.reshape(3, 3, 3), F, T


def _oracle_step(S, alpha, K, B, r_cut, lists):
    rc, Usc, Wsc, PTsc, F1, T1 = H.oracle_forces(S.copy(), "sc", lists=lists, r_cut=r_cut)
    assert rc == 0
    UR, PTR, F2, T2 = H.oracle_forces(S.copy(), "R", alpha=alpha)
    UK, PTK, F3, T3 = H.oracle_forces(S.copy(), "K", alpha=alpha, K=K)
    Ue, _, F4, T4 = H.oracle_forces(S.copy(), "ext", B=B)
    So = S.copy()
    Us = O.lib().o_self_ener(So.c())
    return (np.array([Usc, Wsc, UR, UK, Ue, Us]), np.stack([PTsc, PTR, PTK]), F1 + F2 + F3 + F4,
            T1 + T2 + T3 + T4)


def _cases():
    S, P = H.v251_state()
    R = O.ORun(S, P, seed_init=True)
    v251 = (S.copy(), 0.07, 5, np.array([0.0, 0.0, 5.0]), 10.6, "v251 N=500, one axial dipole")
    R.free()
    S2, P2, inp = H.v25_state()
    R2 = O.ORun(S2, P2, seed_init=True)
    v25 = (S2.copy(), inp["alpha"], inp["K_max"], np.array(inp["B"], dtype=float), inp["sr_cut"],
           "v25 two species, two off-centre dipoles")
    R2.free()
    return [v251, v25]


@pytest.mark.parametrize("case", [0, 1])
def test_adapter_step_against_oracle(probe, case):
    S, alpha, K, B, r_cut, name = _cases()[case]
    _say("case: adapter, %s" % name)
    N = S.N
    rlmin, rlmax = H.list_radii(S, r_cut)
    h, keep = _create(probe, S, alpha, K, B)
    try:
        tot = probe.probe_neigh_list(h, rlmax)
        assert tot >= 0
        off = np.zeros(N + 1, dtype=np.int32)
        nbr = np.zeros(max(tot, 1), dtype=np.int32)
        probe.probe_list(h, off.ctypes.data_as(_ip), nbr.ctypes.data_as(_ip))
        rc, E, PT, F, T = _step(probe, h, N, 0, r_cut, K, alpha)
        assert rc == 0
        # the caller's own list (the oracle's), uploaded by force_sc
        offo, nbro = H.oracle_list(S.copy(), rlmax)
        probe.probe_give_list(h, offo.ctypes.data_as(_ip), np.ascontiguousarray(nbro, np.int32).ctypes.data_as(_ip))
        rc2, E2, PT2, F2, T2 = _step(probe, h, N, 1, r_cut, K, alpha)
        assert rc2 == 0
    finally:
        probe.probe_destroy(h)
    # the list neigh_list filled is the reference's pair set
    assert np.array_equal(off, offo) and np.array_equal(nbr[:tot], nbro)
    Eo, PTo, Fo, To = _oracle_step(S, alpha, K, B, r_cut, (offo, nbro))
    eF, eT = H.max_rel_row(F, Fo), H.max_rel_row(T, To)
    _say("  F %.2e T %.2e  U_sc %.10g U_R %.10g U_K %.10g U_ext %.10g U_self %.10g"
         % (eF, eT, E[0], E[2], E[3], E[4], E[5]))
    assert eF <= 1e-10 and eT <= 1e-10
    for k in (0, 2, 3, 4, 5):
        assert E[k] == pytest.approx(Eo[k], rel=1e-10, abs=1e-12), k
    for j in range(3):
        scale = max(np.abs(PTo[j]).max(), 1e-300)
        assert np.abs(PT[j] - PTo[j]).max() <= 1e-10 * scale, j
    # the same step over the uploaded list
    assert H.max_rel_row(F2, F) <= 1e-12 and H.max_rel_row(T2, T) <= 1e-12
    assert np.allclose(E2, E, rtol=1e-12, atol=1e-14)
