"""Parity of the HIP library (through the C ABI) with the oracle.

The tests share a few long-lived library contexts, each holding several
replicas, and reload the replica they need before each check.

Tolerances (stated per test):
  - neighbour lists: bit-exact pair sets;
  - FP64 path: forces/torques within 1e-11 of the oracle relative to the RMS
    per-particle magnitude (only summation order differs), reciprocal sums
    within 1e-10 (separable phase factors instead of per-k sin/cos);
    energies, virials and pressure tensors within 1e-11 relative;
  - mixed path: relative L2 error of forces and torques <= 1e-5, energies
    <= 1e-6 relative.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

TOL64 = 1e-11
V251, C2, C2CUT = 0, 1, 2   # replicas of the FP64 context
M_C2, M_C2CUT = 0, 1        # replicas of the mixed context


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def msc():
    import mdspherocyl_amd as m
    m.load()
    return m


@pytest.fixture(scope="module")
def v251_init():
    S, P = H.v251_state()
    R = O.ORun(S, P, seed_init=True)
    S0 = S.copy()
    rlmin, rlmax = R.R.r_list_min, R.R.r_list_max
    R.free()
    return S0, P, rlmin, rlmax


@pytest.fixture(scope="module")
def c2_state():
    from mdspherocyl_amd import configs
    desc, st = configs.build("c2")[0]
    return desc, st, H.osys_from_config(desc, st)


def _c2cut_desc(desc):
    from mdspherocyl_amd import replica_desc
    return replica_desc(desc.N, desc.L[0], 3.0 / 3.5, 8, desc.sr_cut, rc=3.5)


@pytest.fixture(scope="module")
def fp64(msc, v251_init, c2_state):
    """One FP64 context: v251 (N=500, all pairs), c2 (N=4096, all pairs),
    c2 with r_c = 3.5."""
    S, P, rlmin, rlmax = v251_init
    desc, st, Sc = c2_state
    _say("context: fp64 [v251, c2, c2 cut-off]")
    sysm = msc.MdSystem([H.v251_desc(), desc, _c2cut_desc(desc)], precision="fp64")
    H.load_osys(sysm, V251, S)
    H.load_osys(sysm, C2, Sc)
    H.load_osys(sysm, C2CUT, Sc)
    yield sysm
    sysm.close()


@pytest.fixture(scope="module")
def mixed(msc, c2_state):
    desc, st, Sc = c2_state
    _say("context: mixed [c2, c2 cut-off]")
    sysm = msc.MdSystem([desc, _c2cut_desc(desc)], precision="mixed")
    H.load_osys(sysm, M_C2, Sc)
    H.load_osys(sysm, M_C2CUT, Sc)
    yield sysm
    sysm.close()


def test_neighbour_list_bit_exact_v251(fp64, v251_init):
    S, P, rlmin, rlmax = v251_init
    _say("case: neighbour list, v251 N=500, r_list=%.6f" % rlmax)
    H.load_osys(fp64, V251, S)
    fp64.neigh_list(V251, rlmax)
    off, nbr = fp64.nblist(V251)
    roff, rnbr = H.oracle_list(S.copy(), rlmax)
    assert np.array_equal(off, roff) and np.array_equal(nbr, rnbr)
    assert nbr.size > 0


def test_neighbour_list_bit_exact_c2(fp64, c2_state):
    desc, st, S = c2_state
    rlmin, rlmax = H.list_radii(S, desc.sr_cut)
    _say("case: neighbour list, c2 N=4096, r_list=%.6f" % rlmax)
    H.load_osys(fp64, C2, S)
    fp64.neigh_list(C2, rlmax)
    off, nbr = fp64.nblist(C2)
    roff, rnbr = H.oracle_list(S.copy(), rlmax)
    assert np.array_equal(off, roff) and np.array_equal(nbr, rnbr)
    assert nbr.size > 100000


@pytest.mark.parametrize("case", ["v251", "c2"])
def test_force_sc_fp64(fp64, v251_init, c2_state, case):
    if case == "v251":
        S, P, rlmin, rlmax = v251_init
        rep, r_cut = V251, 10.6
    else:
        desc, st, S = c2_state
        rlmin, rlmax = H.list_radii(S, desc.sr_cut)
        rep, r_cut = C2, desc.sr_cut
    _say("case: force_sc fp64 %s" % case)
    H.load_osys(fp64, rep, S)
    fp64.neigh_list(rep, rlmax)
    fp64.zero_forces(rep)
    U, W, PT = fp64.force_sc(rep, r_cut)
    F, T = fp64.get_forces(rep)
    So = S.copy()
    lists = H.oracle_list(So, rlmax)
    rc, Uo, Wo, PTo, Fo, To = H.oracle_forces(So, "sc", lists=lists, r_cut=r_cut)
    assert rc == 0
    assert U == pytest.approx(Uo, rel=TOL64, abs=1e-13)
    assert W == pytest.approx(Wo, rel=TOL64, abs=1e-13)
    assert np.allclose(PT, PTo, rtol=TOL64, atol=1e-12 * (1 + np.abs(PTo).max()))
    assert H.max_rel_row(F, Fo) < TOL64
    assert H.max_rel_row(T, To) < TOL64


@pytest.mark.parametrize("case", ["v251", "c2"])
def test_force_dipR_fp64_all_pairs(fp64, v251_init, c2_state, case):
    if case == "v251":
        S, P, rlmin, rlmax = v251_init
        rep, alpha = V251, 0.07
    else:
        desc, st, S = c2_state
        rep, alpha = C2, desc.alpha
    _say("case: force_dipR fp64 all pairs %s" % case)
    H.load_osys(fp64, rep, S)
    fp64.zero_forces(rep)
    U, PT = fp64.force_dipR(rep)
    F, T = fp64.get_forces(rep)
    Uo, PTo, Fo, To = H.oracle_forces(S.copy(), "R", alpha=alpha)
    assert U == pytest.approx(Uo, rel=TOL64)
    assert np.allclose(PT, PTo, rtol=TOL64, atol=TOL64 * np.abs(PTo).max())
    assert H.max_rel_row(F, Fo) < TOL64
    assert H.max_rel_row(T, To) < TOL64


@pytest.mark.parametrize("case", ["v251", "c2"])
def test_force_dipK_fp64(fp64, v251_init, c2_state, case):
    if case == "v251":
        S, P, rlmin, rlmax = v251_init
        rep, alpha, K = V251, 0.07, 5
    else:
        desc, st, S = c2_state
        rep, alpha, K = C2, desc.alpha, desc.K
    _say("case: force_dipK fp64 %s K=%d" % (case, K))
    H.load_osys(fp64, rep, S)
    fp64.zero_forces(rep)
    U, PT = fp64.force_dipK(rep)
    F, T = fp64.get_forces(rep)
    Uo, PTo, Fo, To = H.oracle_forces(S.copy(), "K", alpha=alpha, K=K)
    assert U == pytest.approx(Uo, rel=TOL64)
    assert np.allclose(PT, PTo, rtol=1e-10, atol=1e-10 * np.abs(PTo).max())
    assert H.max_rel_row(F, Fo) < 1e-10
    assert H.max_rel_row(T, To) < 1e-10


def test_ext_field_and_self_energy(fp64, v251_init):
    S, P, rlmin, rlmax = v251_init
    _say("case: ext_field / self_ener v251")
    H.load_osys(fp64, V251, S)
    fp64.zero_forces(V251)
    U = fp64.ext_field(V251)
    F, T = fp64.get_forces(V251)
    Uo, _, Fo, To = H.oracle_forces(S.copy(), "ext", B=(0.0, 0.0, 5.0))
    assert U == pytest.approx(Uo, rel=1e-13)
    assert np.allclose(T, To, rtol=1e-13, atol=1e-14)
    assert fp64.self_ener(V251) == 0.0


def test_cutoff_mode_fp64(fp64, c2_state):
    """Real space restricted to dipole pairs within r_c = 3.5 (pair list built
    with the neighbour list) against the oracle's cut-off sum."""
    desc, st, S = c2_state
    _say("case: force_dipR fp64 cut-off 3.5, c2 state")
    H.load_osys(fp64, C2CUT, S)
    rlmin, rlmax = H.list_radii(S, desc.sr_cut)
    fp64.neigh_list(C2CUT, rlmax)
    fp64.zero_forces(C2CUT)
    U, PT = fp64.force_dipR(C2CUT)
    F, T = fp64.get_forces(C2CUT)
    Uo, PTo, Fo, To = H.oracle_forces(S.copy(), "R", alpha=3.0 / 3.5, rc=3.5)
    assert U == pytest.approx(Uo, rel=TOL64)
    assert H.max_rel_row(F, Fo) < TOL64
    assert H.max_rel_row(T, To) < TOL64


@pytest.mark.parametrize("part", ["R", "K", "Rcut"])
def test_mixed_precision_errors(mixed, c2_state, part):
    """Mixed precision against the FP64 oracle on the c2 state (N=4096):
    relative L2 error of F and T <= 1e-5, the bound of the mixed path.  The
    energy is held to 1e-6 for the real-space parts and to the same 1e-5
    bound for the reciprocal part, whose structure factors run on the matrix
    cores with split fp16 operands (about 2^-21 per product) and whose energy
    is a difference of |S|^2 and the self term (measured: F 4.8e-7, T 3.5e-7,
    U 1.9e-6)."""
    desc, st, S = c2_state
    rep = M_C2CUT if part == "Rcut" else M_C2
    d2 = _c2cut_desc(desc) if part == "Rcut" else desc
    _say("case: mixed precision %s, c2 state" % part)
    H.load_osys(mixed, rep, S)
    if part == "Rcut":
        rlmin, rlmax = H.list_radii(S, desc.sr_cut)
        mixed.neigh_list(rep, rlmax)
    mixed.zero_forces(rep)
    if part == "K":
        U, PT = mixed.force_dipK(rep)
        Uo, PTo, Fo, To = H.oracle_forces(S.copy(), "K", alpha=d2.alpha, K=d2.K)
    else:
        U, PT = mixed.force_dipR(rep)
        Uo, PTo, Fo, To = H.oracle_forces(S.copy(), "R", alpha=d2.alpha, rc=d2.rc)
    F, T = mixed.get_forces(rep)
    eF, eT = H.rel_err(F, Fo), H.rel_err(T, To)
    _say("  rel L2 error F %.3e T %.3e U %.3e" % (eF, eT, abs(U - Uo) / abs(Uo)))
    assert eF <= 1e-5 and eT <= 1e-5
    assert U == pytest.approx(Uo, rel=1e-5 if part == "K" else 1e-6)


def _bring_into_contact(S, rl_short, box):
    """Move one rod of S rigidly towards a neighbour whose segment distance
    lies between rl_short and 2.6 until they are 1.05 apart (inside the WCA
    range 2^(1/6)), keeping it at least 0.9 from every other rod.  Returns
    the moved index or None."""
    from mdspherocyl_amd import configs
    N = S.N
    u = S.ori[:, 2, :]
    for i in range(N):
        others = np.arange(N) != i
        d2 = configs.segdist2(np.repeat(S.r[i:i + 1], N, 0), np.repeat(u[i:i + 1], N, 0), np.full(N, S.len[i]),
                              S.r, u, S.len, box)
        cand = np.nonzero(others & (d2 > rl_short ** 2) & (d2 < 2.6 ** 2))[0]
        for j in cand[:3]:
            v = S.r[j] - S.r[i]
            v -= box * np.round(v / box)
            v /= np.linalg.norm(v)
            lo, hi = 0.0, 3.0
            for _ in range(60):
                mid = 0.5 * (lo + hi)
                r = S.r[i] + mid * v
                dij = configs.segdist2(r[None], u[i:i + 1], S.len[i:i + 1], S.r[j:j + 1], u[j:j + 1],
                                       S.len[j:j + 1], box)[0]
                lo, hi = (mid, hi) if dij > 1.05 ** 2 else (lo, mid)
            r = S.r[i] + lo * v
            dall = configs.segdist2(np.repeat(r[None], N, 0), np.repeat(u[i:i + 1], N, 0), np.full(N, S.len[i]),
                                    S.r, u, S.len, box)
            dall[i] = np.inf
            dall[j] = np.inf
            if dall.min() > 0.9 ** 2 and lo > 0.3:
                S.r[i] = r
                return i
    return None


@pytest.mark.parametrize("moved", [False, True])
def test_force_sc_mixed_contact_rows(mixed, c2_state, moved):
    """The mixed path evaluates WCA (in FP64) over contact rows cut from the
    reference list (segment distance < 2^(1/6) sigma + skin).  With `moved`
    one rod is carried, without a new list, from outside the contact radius
    into WCA contact with a neighbour: force_sc must notice the used-up skin,
    refresh the contact rows and still equal the oracle's force_sc over the
    same reference list (FP64 tolerance)."""
    desc, st, S = c2_state
    rlmin, rlmax = H.list_radii(S, desc.sr_cut)
    _say("case: mixed force_sc over contact rows, c2 state, moved=%s" % moved)
    H.load_osys(mixed, M_C2, S)
    mixed.neigh_list(M_C2, rlmax)
    off, nbr = mixed.nblist(M_C2)
    S2 = S.copy()
    if moved:
        i = _bring_into_contact(S2, 2 ** (1 / 6) + 0.5, desc.L[0])
        assert i is not None
        H.load_osys(mixed, M_C2, S2)
    mixed.zero_forces()
    U, W, PT = mixed.force_sc(M_C2, desc.sr_cut)
    F, T = mixed.get_forces(M_C2)
    rc, Uo, Wo, PTo, Fo, To = H.oracle_forces(S2.copy(), "sc", lists=(off, nbr), r_cut=desc.sr_cut)
    assert rc == 0
    assert U == pytest.approx(Uo, rel=TOL64, abs=1e-13)
    assert W == pytest.approx(Wo, rel=TOL64, abs=1e-13)
    assert H.max_rel_row(F, Fo) < TOL64
    assert H.max_rel_row(T, To) < TOL64


def test_force_sc_overlap_reported(msc):
    """Two rods 0.6 sigma apart: force_sc reports the overlap the reference
    aborts on (v25/functions.cpp:1454)."""
    from mdspherocyl_amd import replica_desc
    N = 2
    _say("case: overlap reported, N=2")
    sysm = msc.MdSystem([replica_desc(N, 20.0, 0.3, 3, 4.5)])
    try:
        r = np.array([[0.0, 0.0, 0.0], [0.6, 0.0, 0.0]])
        ori = np.tile(np.eye(3), (N, 1, 1))
        sysm.set_state(0, r, np.zeros((N, 3)), np.zeros((N, 3)), ori, np.full(N, 3.0), np.ones(N), np.ones(N),
                       np.zeros((N, 1)), np.tile([0, 0, 1.0], (N, 1, 1)), np.ones((N, 1)))
        sysm.neigh_list(0, 6.9)
        sysm.zero_forces()
        with pytest.raises(msc.Overlap):
            sysm.force_sc(0, 4.5)
    finally:
        sysm.close()


def _fixture_lines():
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "v251_run1_output.txt")
    return [[float(x) for x in l.split("|")[1].split()] for l in open(path).read().splitlines()]


def test_step_loop_reproduces_reference_run(msc, v251_init):
    """101 steps of v25/v251/run1 on the GPU (FP64): every printed observable
    of the reference's output.dat within 1e-9 relative."""
    S, P, rlmin, rlmax = v251_init
    _say("case: step loop v251, 101 steps fp64")
    sysm = msc.MdSystem([H.v251_desc()], precision="fp64")
    try:
        H.load_osys(sysm, 0, S)
        sysm.init_run()
        obs = sysm.run(101)
    finally:
        sysm.close()
    ref = _fixture_lines()
    for k, row in enumerate(ref):
        mine = obs[10 * k, 0, :11]
        for j in range(11):
            assert mine[j] == pytest.approx(row[j], rel=1e-9, abs=2e-12), "step %d column %d" % (10 * k, j)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_replicas_match_oracle(msc, precision):
    """Three replicas of different size and real-space mode in one context:
    each replica's total force and torque equal the oracle's for that replica
    alone (FP64: 1e-10 of the RMS per-particle value; mixed: relative L2
    error 1e-5), and 60 steps of the batch reproduce each replica's run in a
    context of its own (1e-12 relative)."""
    from mdspherocyl_amd import configs
    rng = np.random.default_rng(11)
    reps = [configs.make_replica(rng, n, configs.truncated_normal(rng, n, 3.0, 0.3), 0.1, 1.5, 6,
                                 lambda b: 5.6 / b, rc) for n, rc in [(300, 0.0), (517, 3.0), (260, 0.0)]]
    _say("case: batched replicas N=300,517,260 %s" % precision)
    big = msc.MdSystem([d for d, _ in reps], precision=precision)
    try:
        configs.load_into(big, reps)
        big.init_run()
        for k, (d, st) in enumerate(reps):
            S = H.osys_from_config(d, st)
            rlmin, rlmax = H.list_radii(S, d.sr_cut)
            lists = H.oracle_list(S.copy(), rlmax)
            rc, _, _, _, F1, T1 = H.oracle_forces(S.copy(), "sc", lists=lists, r_cut=d.sr_cut)
            _, _, F2, T2 = H.oracle_forces(S.copy(), "R", alpha=d.alpha, rc=d.rc)
            _, _, F3, T3 = H.oracle_forces(S.copy(), "K", alpha=d.alpha, K=d.K)
            F, T = big.get_forces(k)
            if precision == "fp64":
                assert H.max_rel_row(F, F1 + F2 + F3) < 1e-10, "replica %d" % k
                assert H.max_rel_row(T, T1 + T2 + T3) < 1e-10, "replica %d" % k
            else:
                assert H.rel_err(F, F1 + F2 + F3) < 1e-5, "replica %d" % k
                assert H.rel_err(T, T1 + T2 + T3) < 1e-5, "replica %d" % k
        obs = big.run(60)
    finally:
        big.close()
    # the step loop of every replica in the batch equals its run alone
    for k, (d, st) in enumerate(reps):
        one = msc.MdSystem([d], precision=precision)
        try:
            configs.load_into(one, [(d, st)])
            one.init_run()
            alone = one.run(60)
        finally:
            one.close()
        assert np.allclose(obs[:, k, :11], alone[:, 0, :11], rtol=1e-12, atol=1e-14), "replica %d" % k


def test_nve_energy_drift(msc):
    """c1 system with the thermostat switched off (Q = 1e12, gamma = 0, as the
    reference's v25/run1), FP64 and mixed replicas side by side: per-particle
    drift of K+U over 2000 steps below the reference's own 4e-3."""
    from mdspherocyl_amd import configs
    d, st = configs.build("c1")[0]
    d.Q = 1e12
    d.gamma = 0.0
    for precision in ("fp64", "mixed"):
        _say("case: NVE drift c1 %s" % precision)
        sysm = msc.MdSystem([d], precision=precision)
        try:
            configs.load_into(sysm, [(d, st)])
            sysm.init_run()
            obs = sysm.run(2000)
        finally:
            sysm.close()
        E = obs[:, 0, 0] + obs[:, 0, 1]
        drift = float(np.max(np.abs(E - E[0])))
        _say("  max |E - E0| per particle = %.3e" % drift)
        assert drift < 4e-3


def test_list_regions_grow_during_run(msc, monkeypatch):
    """A replica that starts as an aligned lattice at packing fraction 0.3
    melts and its neighbour lists grow.  With regions sized to the initial
    count only (MSC_TIGHT_LISTS=1) a growing rebuild overflows: msc_run rolls
    the chunk back, enlarges the regions and repeats it.  The observables
    equal those of a run with roomy regions to summation-order rounding."""
    from mdspherocyl_amd import configs
    rng = np.random.default_rng(21)
    N = 400
    d, st = configs.make_replica(rng, N, configs.truncated_normal(rng, N, 3.0, 0.3), 0.30, 1.5, 5,
                                 lambda b: 3.0 / 3.5, 3.5)
    runs, repeats = [], []
    for tight in ("1", "0"):
        monkeypatch.setenv("MSC_TIGHT_LISTS", tight)
        _say("case: list regions, N=%d aligned start, tight=%s, 400 steps fp64" % (N, tight))
        sysm = msc.MdSystem([d], precision="fp64")
        try:
            configs.load_into(sysm, [(d, st)])
            sysm.init_run()
            runs.append(sysm.run(400))
            repeats.append(sysm.chunk_repeats())
        finally:
            sysm.close()
    _say("  chunks repeated: tight %d, roomy %d" % tuple(repeats))
    assert repeats[0] >= 1 and repeats[1] == 0
    a, b = runs
    assert np.allclose(a[:, 0, :11], b[:, 0, :11], rtol=1e-8, atol=1e-10)


def _mt_copy(g):
    return np.array(g.mt, dtype=np.uint32), int(g.mti)


def test_cluster_labels_v251_every_call(msc):
    """clust_conf of v25/v251/run1 (min_s = 0.1, every 10 of 101 steps) on the
    GPU: the oracle replays the run and, at each of the 11 calls, the library
    gets the same state and `ran` generator state.  Labels, cluster count and
    the advanced generator state equal the oracle's bit for bit, and the last
    call reproduces the 478 clusters of the reference's conf.dat."""
    import ctypes
    S, P = H.v251_state()
    P.cl_fr = 0  # the clustering calls are made here, in the reference's order
    R = O.ORun(S, P, seed_init=True)
    rlmax = R.R.r_list_max
    sysm = msc.MdSystem([H.v251_desc()], precision="fp64")
    try:
        for step in range(101):
            R.run(1)
            if step % 10:
                continue
            _say("case: cluster labels v251 after step %d" % step)
            H.load_osys(sysm, 0, S)
            sysm.neigh_list(0, rlmax)
            st, mti = _mt_copy(R.R.ran)
            labels, nc, st2, mti2, nasym = sysm.cluster_labels(0, 0.1, st, mti)
            n_ref = O.lib().o_clust_conf(S.c(), 0.1, ctypes.byref(R.R.ran))
            assert nasym == 0
            assert nc == n_ref
            assert np.array_equal(labels, S.clust)
            st_ref, mti_ref = _mt_copy(R.R.ran)
            assert mti2 == mti_ref and np.array_equal(st2, st_ref)
    finally:
        sysm.close()
        R.free()
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "v251_run1_conf.txt")
    lines = open(path).read().splitlines()
    ref = np.array([int(lines[1 + i].split()[-1]) for i in range(500)])
    assert np.array_equal(labels, ref) and nc == 478


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cluster_labels_dense_after_run(msc, precision):
    """A denser system (N=1500, packing fraction 0.25, min_s = 0.6, so that
    clusters of many particles form) after 120 steps of msc_run, which keeps
    the neighbour list: labels and cluster count equal the oracle's
    clust_conf on the state read back, for three successive calls on one
    generator stream."""
    import ctypes
    from mdspherocyl_amd import configs
    rng = np.random.default_rng(31)
    N = 1500
    d, st = configs.make_replica(rng, N, configs.truncated_normal(rng, N, 3.0, 0.3), 0.25, 1.5, 5,
                                 lambda b: 3.0 / 3.5, 3.5)
    _say("case: cluster labels N=%d phi=0.25 after 120 steps %s" % (N, precision))
    sysm = msc.MdSystem([d], precision=precision)
    try:
        configs.load_into(sysm, [(d, st)])
        sysm.init_run()
        sysm.run(120, observe=False)
        g = O.COmt()
        O.lib().omt_seed(ctypes.byref(g), 12345)
        r, p, l, ori = sysm.get_state(0)
        S = H.osys_from_config(d, dict(st, r=r, p=p, l=l, ori=ori))
        sizes = []
        for call in range(3):
            st0, mti = _mt_copy(g)
            labels, nc, st2, mti2, nasym = sysm.cluster_labels(0, 0.6, st0, mti)
            n_ref = O.lib().o_clust_conf(S.c(), 0.6, ctypes.byref(g))
            assert nasym == 0
            assert nc == n_ref
            assert np.array_equal(labels, S.clust), "call %d" % call
            assert (mti2, list(st2)) == (_mt_copy(g)[1], list(_mt_copy(g)[0]))
            sizes.append(np.bincount(labels).max())
        _say("  clusters %d, largest %d" % (nc, sizes[-1]))
        assert sizes[-1] > 5
    finally:
        sysm.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_matches_direct_launches(msc, precision):
    """msc_run replays each step as one captured HIP graph.  Two runs of the
    same two-replica system, one replaying the graph and one launching each
    kernel, over 130 steps from absolute step 40 (three chunks with the
    library's own copies and read-backs between replays, the renormalisation
    at step 100, list rebuilds with a small skin), give bit-identical
    observables and final states."""
    from mdspherocyl_amd import configs
    rng = np.random.default_rng(41)
    N = 400
    d, st = configs.make_replica(rng, N, configs.truncated_normal(rng, N, 3.0, 0.3), 0.15, 1.5, 6,
                                 lambda b: 3.0 / 3.5, 3.5)
    d.dmin = 0.3  # small list skin: rebuilds and contact refreshes within the run
    out = []
    for graph in (1, 0):
        _say("case: graph replay %d, %s, 2 x N=%d, 130 steps" % (graph, precision, N))
        sysm = msc.MdSystem([d, d], precision=precision)
        try:
            configs.load_into(sysm, [(d, st), (d, st)])
            sysm.set_graph(graph)
            sysm.init_run()
            o1 = sysm.run(70, step0=40)
            o2 = sysm.run(60, step0=110)
            out.append((np.concatenate([o1, o2]), sysm.get_state(0), sysm.get_state(1)))
        finally:
            sysm.close()
    (oa, sa0, sa1), (ob, sb0, sb1) = out
    assert int(oa[:, 0, 11].sum()) >= 1
    assert np.array_equal(oa, ob)
    for x, y in zip(sa0 + sa1, sb0 + sb1):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_conf_dat_round_trip_through_device(msc, precision, tmp_path):
    """conf.dat of v25/v251/run1 loaded into a replica built from its
    input.txt (chrg_conf's rules: box line of the file, dipoles rebuilt from
    the body frame), read back from the device and written by
    msc_write_conf: byte-identical to the reference's file.  After 20 steps
    of msc_run the written file reads back to the device state to the 12
    decimals print_conf keeps."""
    from mdspherocyl_amd import refio
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    inp = refio.read_input(os.path.join(gold, "v251_run1_input.txt"))
    conf = os.path.join(gold, "v251_run1_conf.txt")
    c = refio.read_conf(conf, 500, 1)
    _say("case: conf.dat round trip v251 %s" % precision)
    sysm = msc.MdSystem([refio.replica_desc_from_input(inp, box=c["L"])], precision=precision)
    try:
        refio.load_conf(sysm, 0, c, inp)
        out = tmp_path / "conf.dat"
        refio.write_conf(out, refio.conf_from_system(sysm, 0, inp, c["mass"], c["clust"]))
        assert out.read_bytes() == open(conf, "rb").read()
        sysm.init_run()
        sysm.run(20, observe=False)
        refio.write_conf(out, refio.conf_from_system(sysm, 0, inp, c["mass"]))
        c2 = refio.read_conf(out, 500, 1)
        r, p, l, ori = sysm.get_state(0)
        assert not np.array_equal(r, c["r"])
        th = sysm.thermostat(0)
        th = np.array([th[0][0], th[1][0], th[0][1], th[1][1]])
        for a, b in ((r, c2["r"]), (p, c2["p"]), (l, c2["l"]), (ori, c2["ori"]), (th, c2["thermo"])):
            # half a unit of the 12th decimal, plus the parse's own rounding
            assert np.all(np.abs(a - b) <= 5e-13 + 2 * np.spacing(np.abs(a)))
    finally:
        sysm.close()


def test_pair_correl_v251_correlation_dat(msc):
    """acc_correl on the GPU at steps 0, 10, ..., 100 of v25/v251/run1 (the
    oracle replays the run, the library gets each state): the counts channel
    and the dropped count equal the oracle's exactly, the other channels
    within 1e-12 of the sum of their terms' magnitudes (the GPU sums in
    another order), and nor_correl's normalisation of the GPU histograms
    reproduces the reference's correlation.dat to its 8 printed decimals."""
    import gzip
    from mdspherocyl_amd import refio
    S, P = H.v251_state()
    R = O.ORun(S, P, seed_init=True)
    nb = refio.correl_bars(S.L, 0.01)
    ho = np.zeros((1, 4, nb))
    hg = np.zeros((1, 4, nb))
    sysm = msc.MdSystem([H.v251_desc()], precision="fp64")
    try:
        for step in range(101):
            R.run(1)
            if step % 10:
                continue
            _say("case: pair correlations v251 after step %d" % step)
            assert O.lib().o_acc_correl(S.c(), 0.01, nb, O.dptr(ho)) == 0
            H.load_osys(sysm, 0, S)
            _, dropped = sysm.pair_correl(0, 0.01, nb, hist=hg)
            assert dropped == 0
    finally:
        sysm.close()
        R.free()
    assert np.array_equal(hg[:, 0], ho[:, 0])
    npair = ho[0, 0] / 2
    for c, wmax in ((1, 2.0), (2, 8.0), (3, 2.0)):
        assert np.all(np.abs(hg[0, c] - ho[0, c]) <= 1e-12 * wmax * npair + 1e-300)
    V = S.L[0] * S.L[1] * S.L[2]
    cor = 0.0
    for _ in range(11):
        cor += 1. / V
    mine = refio.correlation_text(hg, [500], S.L.min(), 0.01, cor).splitlines()
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "v251_run1_correlation.txt.gz")
    ref = gzip.open(gold, "rt").read().splitlines()
    assert len(mine) == len(ref) == nb
    same = 0
    for a, b in zip(mine, ref):
        assert a.split("&")[0] == b.split("&")[0]
        va = [float(x) for x in a.split("&")[1].replace("|", " ").split()]
        vb = [float(x) for x in b.split("&")[1].replace("|", " ").split()]
        assert np.allclose(va, vb, rtol=0, atol=1.01e-8)
        same += a == b
    _say("  identical lines %d of %d" % (same, len(ref)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pair_correl_three_species(msc, precision):
    """Three species (900, 1100, 500 particles) in a non-cubic box with
    unwrapped positions up to three boxes out: the six histogram sets, the
    dropped pairs beyond the last bar and, with bars of a tenth the width,
    the same histograms split over several LDS bar ranges, against the
    oracle's acc_correl."""
    from mdspherocyl_amd import refio, replica_desc
    rng = np.random.default_rng(17)
    counts = [900, 1100, 500]
    N = sum(counts)
    L = (20.0, 22.0, 18.0)
    S = O.OSys(N, 1, L=L, counts=counts)
    S.r[:] = rng.uniform(-3, 4, size=(N, 3)) * np.array(L)
    for i in range(N):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        S.ori[i] = q
    S.len[:] = 2.0
    S.set_inertia()
    sysm = msc.MdSystem([replica_desc(N, L, 0.5, 2, 5.0, rc=3.0)], precision=precision)
    try:
        H.load_osys(sysm, 0, S)
        for bar_w in (0.05, 0.005):
            nb = refio.correl_bars(L, bar_w)
            _say("case: pair correlations 3 species %s bar_w %g n_bars %d" % (precision, bar_w, nb))
            ho = np.zeros((6, 4, nb))
            d_o = O.lib().o_acc_correl(S.c(), bar_w, nb, O.dptr(ho))
            hg, d_g = sysm.pair_correl(0, bar_w, nb, counts=counts)
            assert d_g == d_o and d_o > 0
            assert np.array_equal(hg[:, 0], ho[:, 0])
            # weight 2 per pair within a species, 1 across: every pair is binned or dropped
            assert hg[:, 0].sum() == N * (N - 1) // 2 - d_o + hg[:3, 0].sum() / 2
            for c, wmax in ((1, 2.0), (2, 8.0), (3, 2.0)):
                assert np.all(np.abs(hg[:, c] - ho[:, c]) <= 1e-12 * wmax * np.maximum(ho[:, 0], 1))
    finally:
        sysm.close()


def test_pair_correl_lattice_ties(msc):
    """A simple cubic lattice (spacing 0.3, 12 per edge, box 12 * 0.3):
    minimum-image quotients on exact halves and distances whose quotient by
    the bar width falls on or next to an integer.  Here a product with the
    reciprocal in place of the reference's divisions picks another image for
    84816 pair components and another bar for 17564 pairs (bars of 0.1) and
    4512 pairs (bars of 0.01), counted on the host below.  Counts channel and
    dropped pairs equal the oracle's."""
    from mdspherocyl_amd import refio, replica_desc
    n, sp = 12, 0.3
    N = n ** 3
    Lb = n * sp
    L = (Lb, Lb, Lb)
    S = O.OSys(N, 1, L=L)
    k = np.arange(N)
    S.r[:] = np.stack([(k % n) * sp, (k // n % n) * sp, (k // (n * n)) * sp], axis=1)
    S.len[:] = 0.1
    S.dia[:] = 0.1
    S.set_inertia()
    # the ties are there: the naive product disagrees with the division
    d = S.r[:, None, :] - S.r[None, :, :]
    assert (np.floor(d / Lb + 0.5) != np.floor(d * (1.0 / Lb) + 0.5)).sum() > 0
    d -= Lb * np.floor(d / Lb + 0.5)
    R = np.sqrt((d * d).sum(-1))
    for bar_w in (0.1, 0.01):
        assert (np.floor(R / bar_w) != np.floor(R * (1.0 / bar_w))).sum() > 0
    sysm = msc.MdSystem([replica_desc(N, L, 0.5, 2, 1.0, rc=0.5)], precision="fp64")
    try:
        H.load_osys(sysm, 0, S)
        for bar_w in (0.1, 0.01):
            nb = refio.correl_bars(L, bar_w)
            _say("case: pair correlations lattice bar_w %g n_bars %d" % (bar_w, nb))
            ho = np.zeros((1, 4, nb))
            d_o = O.lib().o_acc_correl(S.c(), bar_w, nb, O.dptr(ho))
            hg, d_g = sysm.pair_correl(0, bar_w, nb)
            assert d_g == d_o
            assert np.array_equal(hg[:, 0], ho[:, 0])
            assert np.all(np.abs(hg[0, 1:] - ho[0, 1:]) <= 1e-12 * 8 * np.maximum(ho[0, 0], 1))
    finally:
        sysm.close()


def _order_params_fp64(S):
    """G1, G2 and M in FP64 with numpy's symmetric eigensolver, choosing
    among equal magnitudes as oo_params does."""
    u = S.ori[:, 2, :]
    N = len(u)
    Q = (3.0 * u.T @ u - N * np.eye(3)) / (2.0 * N)
    w, v = np.linalg.eigh(Q)
    k = int(np.argmax(np.abs(w)))
    G1 = abs(v[:, k] @ u.sum(0)) / N
    M = (S.dmu[:, :, None] * S.du).sum((0, 1))
    return G1, abs(w[k]), M


def test_order_params_v251(msc):
    """oo_params on the GPU at steps 0, 10, ..., 100 of v25/v251/run1: the
    total moment equals the oracle's within 1e-12 relative; G1 and G2 equal
    an FP64 eigen-decomposition (numpy) within 1e-12 and the oracle's
    single-precision Q tensor and Jacobi rotations (the reference's
    arithmetic) within N * 2^-24, the bound of a serial single-precision sum
    of N terms of size 1/N (the reference's Q is about 6e-6 off here)."""
    import ctypes
    S, P = H.v251_state()
    R = O.ORun(S, P, seed_init=True)
    sysm = msc.MdSystem([H.v251_desc()], precision="mixed")
    try:
        for step in range(101):
            R.run(1)
            if step % 10:
                continue
            _say("case: order parameters v251 after step %d" % step)
            H.load_osys(sysm, 0, S)
            G1, G2, M = sysm.order_params(0)
            g1, g2, Mo = ctypes.c_double(), ctypes.c_double(), np.zeros(3)
            O.lib().o_oo_params(S.c(), ctypes.byref(g1), ctypes.byref(g2), O.dptr(Mo))
            assert np.abs(M - Mo).max() <= 1e-12 * np.abs(Mo).max()
            f1, f2, _ = _order_params_fp64(S)
            assert abs(G1 - f1) <= 1e-12 and abs(G2 - f2) <= 1e-12
            assert abs(G1 - g1.value) <= 500 * 2.0 ** -24 and abs(G2 - g2.value) <= 500 * 2.0 ** -24
    finally:
        sysm.close()
        R.free()


def test_order_params_ref_v251_bit_exact(msc):
    """msc_order_params_ref (the reference's single-precision Q tensor and
    jacobi, sums in particle order) at steps 0, 10, ..., 100 of v25/v251/run1
    equals the oracle's oo_params bit for bit: G1, G2 and M.  The oracle's
    G1/G2 reproduce the G1, G1^2, G2, G2^2 sums of the reference's
    counters.dat exactly (tests/test_oracle_fixtures.py), so the library
    gives the reference's numbers for the same state."""
    import ctypes
    S, P = H.v251_state()
    R = O.ORun(S, P, seed_init=True)
    sysm = msc.MdSystem([H.v251_desc()], precision="fp64")
    try:
        for step in range(101):
            R.run(1)
            if step % 10:
                continue
            _say("case: reference-exact order parameters v251 after step %d" % step)
            H.load_osys(sysm, 0, S)
            G1, G2, M = sysm.order_params_ref(0)
            g1, g2, Mo = ctypes.c_double(), ctypes.c_double(), np.zeros(3)
            O.lib().o_oo_params(S.c(), ctypes.byref(g1), ctypes.byref(g2), O.dptr(Mo))
            _say("  G1 %.17g (oracle %.17g)  G2 %.17g (oracle %.17g)" % (G1, g1.value, G2, g2.value))
            assert G1 == g1.value and G2 == g2.value
            assert np.array_equal(M, Mo)
    finally:
        sysm.close()
        R.free()


def test_order_params_two_dipoles_nematic(msc):
    """N=4096 with two dipoles per particle (moments 2 and -0.5, one off
    the axis) and axes spread about a director (G2 near 0.6): G1, G2 and M
    against the FP64 eigen-decomposition (1e-12) and the oracle (M 1e-12
    relative, G1 and G2 within N * 2^-24, the single-precision Q's bound)."""
    import ctypes
    from mdspherocyl_amd import replica_desc
    rng = np.random.default_rng(23)
    N = 4096
    L = (40.0, 40.0, 40.0)
    S = O.OSys(N, 2, L=L)
    S.r[:] = rng.uniform(0, 40, size=(N, 3))
    n0 = np.array([1.0, 2.0, 2.0]) / 3.0
    for i in range(N):
        u = n0 * np.sign(rng.uniform(-0.3, 1)) + 0.5 * rng.normal(size=3)
        u /= np.linalg.norm(u)
        a = np.cross(u, [0.0, 0.0, 1.0] if abs(u[2]) < 0.9 else [1.0, 0.0, 0.0])
        a /= np.linalg.norm(a)
        S.ori[i] = np.stack([a, np.cross(u, a), u])
    S.len[:] = 2.0
    S.dpc[:, 0], S.dpc[:, 1] = 0.5, -0.5
    S.doc[:, 1] = (0.6, 0.0, 0.8)
    S.dmu[:, 0], S.dmu[:, 1] = 2.0, -0.5
    S.set_inertia()
    S.set_dip()
    sysm = msc.MdSystem([replica_desc(N, L, 0.5, 2, 5.0, rc=3.0, nd=2)], precision="fp64")
    try:
        _say("case: order parameters N=4096 nd=2")
        H.load_osys(sysm, 0, S)
        G1, G2, M = sysm.order_params(0)
        R1, R2, RM = sysm.order_params_ref(0)
    finally:
        sysm.close()
    g1, g2, Mo = ctypes.c_double(), ctypes.c_double(), np.zeros(3)
    O.lib().o_oo_params(S.c(), ctypes.byref(g1), ctypes.byref(g2), O.dptr(Mo))
    _say("  reference-exact: G1 %.17g (oracle %.17g)  G2 %.17g (oracle %.17g)" % (R1, g1.value, R2, g2.value))
    # the single-precision path equals the oracle's bit for bit; M within
    # 1e-15 relative (the library forms the lab-frame dipole axes of the
    # tilted slot from the body frame, the oracle stores them)
    assert R1 == g1.value and R2 == g2.value
    assert np.abs(RM - Mo).max() <= 1e-15 * np.abs(Mo).max()
    f1, f2, Mf = _order_params_fp64(S)
    assert 0.3 < G2 < 0.9 and 0.1 < G1 < 0.9
    assert abs(G1 - f1) <= 1e-12 and abs(G2 - f2) <= 1e-12
    assert np.abs(M - Mo).max() <= 1e-12 * np.abs(Mo).max() and np.abs(M - Mf).max() <= 1e-12 * np.abs(Mf).max()
    assert abs(G1 - g1.value) <= N * 2.0 ** -24 and abs(G2 - g2.value) <= N * 2.0 ** -24


def test_reference_run_from_library_initial_state(msc):
    """v25/v251/run1 end to end without the oracle: the library's Signal=0
    initial state (msc_ini_state: lattice, 1000 GCAmix moves, ini_mom from
    the generator seeded 1), loaded with the run's dipoles, then 101 steps of
    msc_run (FP64): every printed observable of the reference's output.dat
    within 1e-9 relative."""
    import math
    L = math.pow(500.0 / ((1.0 * 1.0) * 0.0013), 1.0 / 3)
    st = msc.ini_state(500, (L, L, L), 9.0, 1.0, 1.0, 1.0)
    _say("case: v251 run from msc_ini_state, 101 steps fp64")
    sysm = msc.MdSystem([H.v251_desc()], precision="fp64")
    try:
        N = 500
        sysm.set_state(0, st["r"], st["p"], st["l"], st["ori"], np.full(N, 9.0), np.ones(N), np.ones(N),
                       np.zeros((N, 1)), np.tile([0.0, 0.0, 1.0], (N, 1, 1)), np.full((N, 1), 4.481))
        sysm.init_run()
        obs = sysm.run(101)
    finally:
        sysm.close()
    ref = _fixture_lines()
    for k, row in enumerate(ref):
        mine = obs[10 * k, 0, :11]
        for j in range(11):
            assert mine[j] == pytest.approx(row[j], rel=1e-9, abs=2e-12), "step %d column %d" % (10 * k, j)


def test_fp64_contact_rows_match_reference_rows(msc):
    """An FP64 context with msc_set_contact_rows(1) (WCA over the contact
    rows, refreshed as the particles move) against the default (every
    reference-row pair): 101 steps of v25/v251/run1 give the same printed
    observables to 1e-12 and the reference's output.dat to 1e-9; and a
    denser system (N=1500, packing fraction 0.25) after 200 steps has the
    same state to 1e-12."""
    from mdspherocyl_amd import configs
    import math
    L = math.pow(500.0 / ((1.0 * 1.0) * 0.0013), 1.0 / 3)
    st = msc.ini_state(500, (L, L, L), 9.0, 1.0, 1.0, 1.0)
    N = 500
    args = (st["r"], st["p"], st["l"], st["ori"], np.full(N, 9.0), np.ones(N), np.ones(N), np.zeros((N, 1)),
            np.tile([0.0, 0.0, 1.0], (N, 1, 1)), np.full((N, 1), 4.481))
    obs = []
    for contact in (False, True):
        _say("case: v251 fp64 contact rows %s, 101 steps" % contact)
        sysm = msc.MdSystem([H.v251_desc()], precision="fp64")
        try:
            sysm.set_contact_rows(contact)
            sysm.set_state(0, *args)
            sysm.init_run()
            obs.append(sysm.run(101))
        finally:
            sysm.close()
    assert np.allclose(obs[1][:, 0, :11], obs[0][:, 0, :11], rtol=1e-12, atol=1e-14)
    ref = _fixture_lines()
    for k, row in enumerate(ref):
        for j in range(11):
            assert obs[1][10 * k, 0, j] == pytest.approx(row[j], rel=1e-9, abs=2e-12)
    rng = np.random.default_rng(41)
    n = 1500
    d, s0 = configs.make_replica(rng, n, configs.truncated_normal(rng, n, 3.0, 0.3), 0.25, 1.5, 5,
                                 lambda b: 3.0 / 3.5, 3.5)
    states = []
    for contact in (False, True):
        _say("case: N=1500 phi=0.25 fp64 contact rows %s, 200 steps" % contact)
        sysm = msc.MdSystem([d], precision="fp64")
        try:
            sysm.set_contact_rows(contact)
            configs.load_into(sysm, [(d, s0)])
            sysm.init_run()
            sysm.run(200, observe=False)
            states.append(sysm.get_state(0))
        finally:
            sysm.close()
    for a, b in zip(*states):
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(a).max())


def test_fp64_contact_rows_switched_on_mid_run(msc):
    """Contact rows switched on between two msc_run calls of an FP64
    context (they are refreshed where the particles have used up their skin
    since the last list build): 100 + 100 steps equal 200 steps on the
    reference rows to 1e-12."""
    from mdspherocyl_amd import configs
    rng = np.random.default_rng(43)
    n = 1500
    d, s0 = configs.make_replica(rng, n, configs.truncated_normal(rng, n, 3.0, 0.3), 0.25, 1.5, 5,
                                 lambda b: 3.0 / 3.5, 3.5)
    states = []
    for switch in (False, True):
        _say("case: N=1500 fp64, contact rows switched on at step 100: %s" % switch)
        sysm = msc.MdSystem([d], precision="fp64")
        try:
            configs.load_into(sysm, [(d, s0)])
            sysm.init_run()
            sysm.run(100, observe=False)
            if switch:
                sysm.set_contact_rows(True)
            sysm.run(100, step0=100, observe=False)
            states.append(sysm.get_state(0))
        finally:
            sysm.close()
    for a, b in zip(*states):
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(a).max())
