"""The benchmark workloads at their full sizes on the GPU (through the C ABI).

  c3  N=32768 mixed and FP64 against the oracle at the initial state, per
      particle; 5000 thermostatted steps, then 5000 steps with the
      thermostat off (Q = 1e12, gamma = 0, as v25/run1/input.txt:12-13) and
      the energy drift of both precisions.
  c4  64 replicas of 512..8192 rods in one context: every replica's mixed
      forces against its own FP64 evaluation, FP64 replicas in the batch
      against single-replica contexts, and 200 steps of both.
  c5  N=262144: mixed against FP64 per particle; cluster labels at two
      frames (the initial state and after 100 steps), geometric and with
      the dipolar clause (pair energy below -1 kT), checked against contact
      pairs found independently (a k-d tree over the centres and the
      oracle's dist_test) and the creation-order numbering replayed here.

Tolerances, stated per check:
  - per-particle errors are max_i |a_i - b_i| / rms_j |b_j| (max_rel_rms);
    the worst particle against its own magnitude (max_rel_own) is printed;
  - FP64 against the oracle: 1e-11; mixed against FP64 or the oracle: 1e-5
    for forces, torques and energies (relative), the bound the workloads
    state; where a total energy sums parts of opposite sign (c4's dilute
    replicas), its error is taken against the parts' magnitudes, and per
    particle it stays below 1e-6 kT.  The measured figures are printed.
"""
import ctypes
import math
import time

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _fmt(e):
    return "F max/rms %.2e max/own %.2e rel L2 %.2e | T max/rms %.2e max/own %.2e rel L2 %.2e | U %.2e" % (
        e["F"]["max_rel_rms"], e["F"]["max_rel_own"], e["F"]["rel_l2"], e["T"]["max_rel_rms"],
        e["T"]["max_rel_own"], e["T"]["rel_l2"], e["U"])


def force_split(sysm, rep, sr_cut):
    sysm.zero_forces(rep)
    Usc, _, _ = sysm.force_sc(rep, sr_cut)
    UR, _ = sysm.force_dipR(rep)
    UK, _ = sysm.force_dipK(rep)
    F, T = sysm.get_forces(rep)
    return F, T, Usc + UR + UK, (Usc, UR, UK)


@pytest.fixture(scope="module")
def msc():
    import mdspherocyl_amd as m
    m.load()
    return m


@pytest.fixture(scope="module")
def c3():
    from mdspherocyl_amd import configs
    t0 = time.time()
    desc, st = configs.build("c3")[0]
    _say("c3 built in %.1f s" % (time.time() - t0))
    return desc, st


def test_c3_initial_forces_per_particle(msc, c3):
    from mdspherocyl_amd import accuracy, configs
    desc, st = c3
    out = {}
    lists = None
    for prec in ("fp64", "mixed"):
        _say("case: c3 N=%d %s initial forces" % (desc.N, prec))
        s = msc.MdSystem([desc], precision=prec)
        try:
            configs.load_into(s, [(desc, st)])
            s.init_run()
            out[prec] = force_split(s, 0, desc.sr_cut)
            if prec == "fp64":
                lists = s.nblist(0)
        finally:
            s.close()
    _say("case: c3 oracle (serial) on the same state and list")
    S = H.osys_from_config(desc, st)
    rc, Usc, _, _, F1, T1 = H.oracle_forces(S.copy(), "sc", lists=lists, r_cut=desc.sr_cut)
    assert rc == 0
    UR, _, F2, T2 = H.oracle_forces(S.copy(), "R", alpha=desc.alpha, rc=desc.rc)
    UK, _, F3, T3 = H.oracle_forces(S.copy(), "K", alpha=desc.alpha, K=desc.K)
    Fo, To, Uo = F1 + F2 + F3, T1 + T2 + T3, Usc + UR + UK
    e64 = accuracy.force_errors(out["fp64"][0], out["fp64"][1], out["fp64"][2], Fo, To, Uo)
    emo = accuracy.force_errors(out["mixed"][0], out["mixed"][1], out["mixed"][2], Fo, To, Uo)
    emd = accuracy.force_errors(*out["mixed"][:3], *out["fp64"][:3])
    _say("  FP64  vs oracle: " + _fmt(e64))
    _say("  mixed vs oracle: " + _fmt(emo))
    _say("  mixed vs FP64  : " + _fmt(emd))
    for part, a, b in zip(("WCA", "real", "recip"), out["mixed"][3], (Usc, UR, UK)):
        _say("  mixed %s energy %.12g oracle %.12g rel %.2e" % (part, a, b, abs(a - b) / abs(b) if b else 0.0))
    assert e64["F"]["max_rel_rms"] <= 1e-11 and e64["T"]["max_rel_rms"] <= 1e-11 and e64["U"] <= 1e-11
    for e in (emo, emd):
        assert e["F"]["max_rel_rms"] <= 1e-5 and e["T"]["max_rel_rms"] <= 1e-5
        assert e["U"] <= 1e-5
    for a, b in zip(out["mixed"][3], (Usc, UR, UK)):
        assert abs(a - b) <= 1e-5 * abs(b) + 1e-9 * abs(Uo)


def test_c3_thermostat_then_constant_energy_drift(msc, c3):
    """5000 steps with the Nose thermostat of the workload, then 5000 steps
    from that state with the thermostat switched off (Q = 1e12, gamma = 0,
    s = 0), both precisions.  The drift is max_t |E(t) - E(0)| per particle
    with E = K + U (print_out's columns).  Bounds: the mixed drift stays
    within twice the FP64 drift plus 1e-4 (one part in 10^4 of E per
    particle), and neither exceeds 4e-3, the reference's own drift over its
    constant-energy run (v25/run1/output.dat)."""
    from mdspherocyl_amd import configs, replica_desc
    desc, st = c3
    drift = {}
    for prec in ("fp64", "mixed"):
        _say("case: c3 %s 5000 thermostatted steps" % prec)
        s = msc.MdSystem([desc], precision=prec)
        try:
            configs.load_into(s, [(desc, st)])
            s.init_run()
            obs = s.run(5000)
            r, p, l, ori = s.get_state(0)
        finally:
            s.close()
        assert np.all(np.isfinite(obs))
        nve = replica_desc(desc.N, desc.L[0], desc.alpha, desc.K, desc.sr_cut, rc=desc.rc, temp=desc.temp,
                           dt=desc.dt, Q=1e12, gamma=0.0, s0=0.0, dens=desc.dens)
        _say("case: c3 %s 5000 steps, thermostat off" % prec)
        s = msc.MdSystem([nve], precision=prec)
        try:
            s.set_state(0, r, p, l, ori, st["length"], st["dia"], st["mass"], st["dpc"], st["doc"], st["dmu"])
            s.set_thermostat(0, np.zeros(2), np.zeros(2))
            s.init_run()
            obs2 = s.run(5000)
        finally:
            s.close()
        E = obs2[:, 0, 0] + obs2[:, 0, 1]
        drift[prec] = float(np.max(np.abs(E - E[0])))
        _say("  %s: T after the thermostatted run %.4f; constant-energy E per particle %.6f, drift %.3e, "
             "rebuilds %d" % (prec, obs[-1, 0, 0] / 3, E[0], drift[prec], int(obs2[:, 0, 11].sum())))
    assert drift["fp64"] <= 4e-3 and drift["mixed"] <= 4e-3
    assert drift["mixed"] <= 2 * drift["fp64"] + 1e-4


def test_c4_every_replica_against_its_own_fp64(msc):
    from mdspherocyl_amd import accuracy, configs
    t0 = time.time()
    reps = configs.build("c4")
    descs = [d for d, _ in reps]
    _say("case: c4 %d replicas, sum N %d, built in %.1f s" % (len(reps), sum(d.N for d in descs),
                                                            time.time() - t0))
    mix = msc.MdSystem(descs, precision="mixed")
    f64 = msc.MdSystem(descs, precision="fp64")
    try:
        configs.load_into(mix, reps)
        configs.load_into(f64, reps)
        mix.init_run()
        f64.init_run()
        worst = {"F": 0.0, "T": 0.0, "U": 0.0, "Uparts": 0.0, "UkT": 0.0}
        own = {"F": 0.0, "T": 0.0}
        ref = {}
        for rp, d in enumerate(descs):
            Fm, Tm, Um, _ = force_split(mix, rp, d.sr_cut)
            Fd, Td, Ud, parts = force_split(f64, rp, d.sr_cut)
            ref[rp] = (Fd, Td, Ud)
            e = accuracy.force_errors(Fm, Tm, Um, Fd, Td, Ud)
            for k in ("F", "T"):
                worst[k] = max(worst[k], e[k]["max_rel_rms"])
                own[k] = max(own[k], e[k]["max_rel_own"])
            # the total sums parts of opposite sign (real space < 0 < WCA,
            # reciprocal); its error is measured against their magnitudes
            # and per particle in units of kT, the plain ratio is printed
            eu = abs(Um - Ud) / sum(abs(x) for x in parts)
            ukt = abs(Um - Ud) / d.N / d.temp
            worst["U"] = max(worst["U"], e["U"])
            worst["Uparts"] = max(worst["Uparts"], eu)
            worst["UkT"] = max(worst["UkT"], ukt)
            assert e["F"]["max_rel_rms"] <= 1e-5 and e["T"]["max_rel_rms"] <= 1e-5, (rp, e)
            assert eu <= 1e-5 and ukt <= 1e-6, (rp, eu, ukt, parts)
        _say("  worst replica, mixed vs FP64: F max/rms %.2e (max/own %.2e), T max/rms %.2e (max/own %.2e), "
             "U %.2e of |U_sc|+|U_R|+|U_K| (%.2e kT per particle; plain relative %.2e)"
             % (worst["F"], own["F"], worst["T"], own["T"], worst["Uparts"], worst["UkT"], worst["U"]))
        # FP64 replicas in the batch equal single-replica contexts
        for rp in (0, 31, 63):
            one = msc.MdSystem([descs[rp]], precision="fp64")
            try:
                configs.load_into(one, [reps[rp]])
                one.init_run()
                F1, T1, U1, _ = force_split(one, 0, descs[rp].sr_cut)
            finally:
                one.close()
            Fd, Td, Ud = ref[rp]
            e = accuracy.force_errors(F1, T1, U1, Fd, Td, Ud)
            _say("  replica %d (N=%d) batched FP64 vs alone: F %.2e T %.2e U %.2e"
                 % (rp, descs[rp].N, e["F"]["max_rel_rms"], e["T"]["max_rel_rms"], e["U"]))
            assert e["F"]["max_rel_rms"] <= 1e-13 and e["T"]["max_rel_rms"] <= 1e-13 and e["U"] <= 1e-13
        _say("case: c4 200 steps, mixed and FP64")
        om = mix.run(200)
        od = f64.run(200)
        assert np.all(np.isfinite(om)) and np.all(np.isfinite(od))
        # K and U are per particle (print_out's columns): differences in kT
        temps = np.array([d.temp for d in descs])
        dK = np.abs(om[:, :, 0] - od[:, :, 0]) / temps
        dU = np.abs(om[:, :, 1] - od[:, :, 1]) / temps
        _say("  200 steps, mixed vs FP64 per replica, per particle in kT: K max %.2e (step 1: %.2e), U max %.2e "
             "(step 1: %.2e); rebuilds mixed %d FP64 %d"
             % (dK.max(), dK[0].max(), dU.max(), dU[0].max(), int(om[:, :, 11].sum()), int(od[:, :, 11].sum())))
        assert dK[0].max() <= 1e-6 and dU[0].max() <= 1e-6
        assert dK.max() <= 1e-3 and dU.max() <= 1e-3
    finally:
        mix.close()
        f64.close()


# ---------------------------------------------------------------------------
# c5
# ---------------------------------------------------------------------------
def contact_partition(S, box, reach, min_s, u_cut):
    """Components of the contact graph (dist_test in either order, and the
    dipolar clause when u_cut is finite): candidate pairs from a periodic
    k-d tree over the centres, contacts from the oracle."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    from scipy.spatial import cKDTree
    x = np.mod(S.r, box)
    x[x >= box] = 0.0
    pairs = cKDTree(x, boxsize=box).query_pairs(reach, output_type="ndarray")
    a = np.ascontiguousarray(pairs[:, 0], dtype=np.int32)
    b = np.ascontiguousarray(pairs[:, 1], dtype=np.int32)
    hit = np.zeros(len(a), dtype=np.uint8)
    O.lib().o_contact_pairs(S.c(), len(a), O.iptr(a), O.iptr(b), min_s, u_cut,
                            hit.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
    hit = hit.astype(bool)
    g = coo_matrix((np.ones(hit.sum()), (a[hit], b[hit])), shape=(S.N, S.N))
    n, comp = connected_components(g, directed=False)
    return n, comp, len(a), int(hit.sum())


def replay_numbering(comp, mt_state, mt_index):
    """clust_conf's numbering (v25/functions.cpp:2950-2970, 3282-3324): one
    uniform u per cluster, the seed is the floor(u * left)-th particle not yet
    in a cluster (ascending index), and the seed's component gets the next
    number.  A Fenwick tree over the remaining particles finds each seed."""
    N = len(comp)
    g = O.COmt()
    for i in range(624):
        g.mt[i] = int(mt_state[i])
    g.mti = int(mt_index)
    tree = [0] * (N + 1)
    for i in range(1, N + 1):
        tree[i] += 1
        j = i + (i & -i)
        if j <= N:
            tree[j] += tree[i]
    order = np.argsort(comp, kind="stable")
    bounds = np.searchsorted(comp[order], np.arange(comp.max() + 2))
    top = 1
    while top * 2 <= N:
        top *= 2
    labels = np.full(N, -1, dtype=np.int64)
    left, nc = N, 0
    while left > 0:
        ip = int(math.floor(O.lib().omt_uniform(ctypes.byref(g)) * left))
        pos = 0
        step = top
        while step:
            if pos + step <= N and tree[pos + step] <= ip:
                pos += step
                ip -= tree[pos]
            step >>= 1
        seed = pos  # 0-based index of the (ip+1)-th remaining particle
        c = comp[seed]
        members = order[bounds[c]:bounds[c + 1]]
        labels[members] = nc
        for m in members.tolist():
            i = m + 1
            while i <= N:
                tree[i] -= 1
                i += i & -i
        left -= len(members)
        nc += 1
    return labels, nc, g


def test_c5_mixed_against_fp64_and_cluster_labels(msc):
    from mdspherocyl_amd import accuracy, configs
    t0 = time.time()
    desc, st = configs.build("c5")[0]
    _say("case: c5 N=%d built in %.1f s (box %.2f)" % (desc.N, time.time() - t0, desc.L[0]))
    mix = msc.MdSystem([desc], precision="mixed")
    f64 = msc.MdSystem([desc], precision="fp64")
    try:
        configs.load_into(mix, [(desc, st)])
        configs.load_into(f64, [(desc, st)])
        mix.init_run()
        f64.init_run()
        Fm, Tm, Um, _ = force_split(mix, 0, desc.sr_cut)
        Fd, Td, Ud, _ = force_split(f64, 0, desc.sr_cut)
        e = accuracy.force_errors(Fm, Tm, Um, Fd, Td, Ud)
        _say("  c5 mixed vs FP64: " + _fmt(e))
        assert e["F"]["max_rel_rms"] <= 1e-5 and e["T"]["max_rel_rms"] <= 1e-5 and e["U"] <= 1e-5
        f64.close()
        S = H.osys_from_config(desc, st)
        reach = float(st["length"].max()) + 1.0 + 0.2 + 1e-6
        mt, mti = msc.mt19937_state(5)
        for frame in (0, 100):
            if frame:
                _say("case: c5 mixed 100 steps")
                mix.run(100, observe=False)
                r, p, l, ori = mix.get_state(0)
                S.r[:], S.ori[:] = r, ori
                S.set_dip()
            for u_cut in (math.inf, -1.0 * desc.temp):
                tag = "geometric" if math.isinf(u_cut) else "dipolar (U_dd < -1 kT)"
                _say("case: c5 clusters at step %d, %s" % (frame, tag))
                t1 = time.time()
                labels, nclust, mt2, mti2, nasym = mix.cluster_labels(0, 0.2, mt, mti,
                                                                      u_cut=None if math.isinf(u_cut) else u_cut)
                t_gpu = time.time() - t1
                n, comp, ncand, ncont = contact_partition(S, desc.L[0], reach, 0.2, u_cut)
                ref, nref, g = replay_numbering(comp, mt, mti)
                _say("  %d clusters (largest %d), %d contacts among %d candidate pairs, asymmetric %d; "
                     "library call %.3f s" % (nclust, np.bincount(labels).max(), ncont, ncand, nasym, t_gpu))
                assert nclust == n == nref
                assert np.array_equal(labels, ref)
                assert list(mt2) == list(g.mt) and mti2 == g.mti
                mt, mti = mt2, mti2
    finally:
        mix.close()
        f64.close()
