"""Several dipole slots per particle on the GPU (through the C ABI).

The system of v25/input.txt: two species of 50 rods (L = D = 1), each with
two off-centre dipoles (at -0.5 and +0.5 along the axis, tilted in the body
frame, moments 1.5 and 2), K = 10, alpha = 0.75, NHL thermostat gamma = 1,
dt = 1e-4, from its Signal=0 start made by the oracle (ini_pos, 1000 GCAmix
moves, ini_mom).  It exercises the reciprocal sum's intra-particle exclusion
over two slots (v25/functions.cpp:1780-1925), the lever-arm torques of the
real-space and reciprocal sums and the intra-particle self energy.

Tolerances:
  - FP64 force routines at the start: forces and torques within 1e-11 of the
    oracle per particle relative to the rms magnitude (pairs), 1e-10 for the
    reciprocal sum (separable phase factors instead of per-k sin/cos);
    energies and pressure tensors within 1e-11 / 1e-10 relative;
  - mixed contexts: per particle within 1e-5 of the rms magnitude, energies
    1e-6 relative (their reciprocal sum is FP64 as well);
  - the step loop: all 661 steps printed at v25/output.dat lines 1220-1880
    (K U P Uss Pss UdipR PdipR UdipK PdipK |P|^2 |L|^2) within 1e-9 relative.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def msc():
    import mdspherocyl_amd as m
    m.load()
    return m


@pytest.fixture(scope="module")
def v25_start():
    S, P, inp = H.v25_state()
    R = O.ORun(S, P, seed_init=True)
    S0 = S.copy()
    R.free()
    return S0, P, inp


def _pt(PT):
    return np.asarray(PT, dtype=np.float64).reshape(3, 3)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_force_routines_two_slots(msc, v25_start, precision):
    S, P, inp = v25_start
    _say("case: v25 two slots, N=%d, force routines %s" % (S.N, precision))
    rlmin, rlmax = H.list_radii(S, inp["sr_cut"])
    sysm = msc.MdSystem([H.v25_desc(S, inp)], precision=precision)
    try:
        H.load_osys(sysm, 0, S)
        sysm.neigh_list(0, rlmax)
        out = {}
        for which in ("sc", "R", "K"):
            sysm.zero_forces()
            if which == "sc":
                U, W, PT = sysm.force_sc(0, inp["sr_cut"])
            elif which == "R":
                U, PT = sysm.force_dipR(0)
            else:
                U, PT = sysm.force_dipK(0)
            F, T = sysm.get_forces(0)
            out[which] = (U, _pt(PT), F, T)
        Uself = sysm.self_ener(0)
    finally:
        sysm.close()
    lists = H.oracle_list(S.copy(), rlmax)
    ref = {}
    rc, U, W, PT, F, T = H.oracle_forces(S.copy(), "sc", lists=lists, r_cut=inp["sr_cut"])
    assert rc == 0
    ref["sc"] = (U, PT, F, T)
    ref["R"] = H.oracle_forces(S.copy(), "R", alpha=inp["alpha"])
    ref["K"] = H.oracle_forces(S.copy(), "K", alpha=inp["alpha"], K=inp["K_max"])
    fp64 = precision == "fp64"
    for which in ("sc", "R", "K"):
        U, PT, F, T = out[which]
        Uo, PTo, Fo, To = ref[which]
        eF, eT = H.max_rel_row(F, Fo), H.max_rel_row(T, To)
        eU = abs(U - Uo) / max(abs(Uo), 1e-300)
        ePT = float(np.max(np.abs(PT - PTo)) / max(np.max(np.abs(PTo)), 1e-300))
        _say("  %s: U %.12g (oracle %.12g) F %.2e T %.2e U %.2e PT %.2e" % (which, U, Uo, eF, eT, eU, ePT))
        tf = (1e-10 if which == "K" else 1e-11) if fp64 else 1e-5
        tu = (1e-10 if which == "K" else 1e-11) if fp64 else 1e-6
        if which == "sc" and not np.any(Fo):
            assert not np.any(F) and U == 0.0
            continue
        assert eF <= tf and eT <= tf, which
        assert eU <= tu and ePT <= tu, which
    # the intra-particle dipole energy (self_ener, v25/functions.cpp:2027-2069)
    So = S.copy()  # kept alive: its arrays back the struct the call reads
    Uo_self = O.lib().o_self_ener(So.c())
    _say("  self_ener %.12g (oracle %.12g)" % (Uself, Uo_self))
    assert Uself == pytest.approx(Uo_self, rel=1e-12)


def test_reciprocal_single_particle_is_self_free(msc):
    """One particle with two slots alone in the box: the reciprocal energy
    and force vanish (every k term is the particle's own, removed by the
    exclusion); its torque is the oracle's (the exclusion leaves no torque
    either: within 1e-12 of the largest slot-pair term)."""
    S = O.OSys(1, 2, L=(9.0, 9.0, 9.0))
    S.r[0] = (0.3, -1.2, 2.0)
    S.len[:] = 2.0
    S.mass[:] = 1.0
    S.dpc[0] = (-0.7, 0.6)
    S.doc[0, 0] = (0.6, 0.0, 0.8)
    S.doc[0, 1] = (0.0, 1.0, 0.0)
    S.dmu[0] = (1.5, 2.0)
    c, s_ = np.cos(0.4), np.sin(0.4)
    S.ori[0] = np.array([[c, s_, 0.0], [-s_, c, 0.0], [0.0, 0.0, 1.0]]) @ np.array(
        [[1.0, 0.0, 0.0], [0.0, np.cos(0.9), np.sin(0.9)], [0.0, -np.sin(0.9), np.cos(0.9)]])
    S.set_inertia()
    S.set_dip()
    from mdspherocyl_amd import replica_desc
    sysm = msc.MdSystem([replica_desc(1, 9.0, 0.8, 6, 2.5, nd=2)], precision="fp64")
    try:
        H.load_osys(sysm, 0, S)
        sysm.zero_forces()
        U, PT = sysm.force_dipK(0)
        F, T = sysm.get_forces(0)
    finally:
        sysm.close()
    Uo, PTo, Fo, To = H.oracle_forces(S.copy(), "K", alpha=0.8, K=6)
    _say("case: one particle two slots: U %.3e (oracle %.3e) |F| %.3e |T| %.3e (oracle %.3e)"
         % (U, Uo, np.abs(F).max(), np.abs(T).max(), np.abs(To).max()))
    assert abs(U) < 1e-12 and abs(Uo) < 1e-12
    assert np.abs(F).max() < 1e-12
    assert np.allclose(T, To, rtol=0, atol=1e-12)


def _v25_lines():
    lines = open(os.path.join(GOLD, "v25_output_seg.txt")).read().splitlines()
    return [[float(x) for x in ln.split("|")[1].split()] for ln in lines]


def test_step_loop_reproduces_two_species_two_slot_run(msc, v25_start):
    """661 steps of v25/input.txt through msc_run (FP64) from the Signal=0
    start: every printed observable of v25/output.dat lines 1220-1880 within
    1e-9 relative."""
    S, P, inp = v25_start
    _say("case: v25 two species two slots, 661 steps fp64")
    sysm = msc.MdSystem([H.v25_desc(S, inp)], precision="fp64")
    try:
        H.load_osys(sysm, 0, S)
        sysm.init_run()
        obs = sysm.run(661)
    finally:
        sysm.close()
    ref = _v25_lines()
    assert len(ref) == 661
    worst = 0.0
    for k, row in enumerate(ref):
        mine = obs[k, 0, :11]
        for j in range(11):
            d = abs(mine[j] - row[j]) / max(abs(row[j]), 1e-300)
            if abs(mine[j] - row[j]) > 2e-12:
                worst = max(worst, d)
            assert mine[j] == pytest.approx(row[j], rel=1e-9, abs=2e-12), "step %d column %d" % (k + 1, j)
    _say("  661 steps: worst relative deviation %.2e" % worst)
