"""Parity of the HIP library (through the C ABI) with the oracle.

Tolerances (stated per test):
  - neighbour lists: bit-exact pair sets;
  - FP64 path: forces/torques within 1e-11 of the oracle relative to the RMS
    per-particle magnitude (only summation order differs); energies, virials
    and pressure tensors within 1e-11 relative;
  - mixed path: relative L2 error of forces and torques <= 1e-5, energies
    <= 1e-6 relative.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

TOL64 = 1e-11


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


def _system(msc, desc, precision="fp64"):
    return msc.MdSystem([desc], precision=precision)


def test_neighbour_list_bit_exact_v251(msc, v251_init):
    S, P, rlmin, rlmax = v251_init
    _say("case: neighbour list, v251 N=500, r_list=%.6f" % rlmax)
    sysm = _system(msc, H.v251_desc())
    H.load_osys(sysm, 0, S)
    sysm.neigh_list(0, rlmax)
    off, nbr = sysm.nblist(0)
    roff, rnbr = H.oracle_list(S.copy(), rlmax)
    assert np.array_equal(off, roff) and np.array_equal(nbr, rnbr)
    assert nbr.size > 0


def test_neighbour_list_bit_exact_c2(msc, c2_state):
    desc, st, S = c2_state
    rlmin, rlmax = H.list_radii(S, desc.sr_cut)
    _say("case: neighbour list, c2 N=4096, r_list=%.6f" % rlmax)
    sysm = _system(msc, desc)
    H.load_osys(sysm, 0, S)
    sysm.neigh_list(0, rlmax)
    off, nbr = sysm.nblist(0)
    roff, rnbr = H.oracle_list(S.copy(), rlmax)
    assert np.array_equal(off, roff) and np.array_equal(nbr, rnbr)
    assert nbr.size > 100000


@pytest.mark.parametrize("case", ["v251", "c2"])
def test_force_sc_fp64(msc, v251_init, c2_state, case):
    if case == "v251":
        S, P, rlmin, rlmax = v251_init
        desc, r_cut = H.v251_desc(), 10.6
    else:
        desc, st, S = c2_state
        rlmin, rlmax = H.list_radii(S, desc.sr_cut)
        r_cut = desc.sr_cut
    _say("case: force_sc fp64 %s" % case)
    sysm = _system(msc, desc)
    H.load_osys(sysm, 0, S)
    sysm.neigh_list(0, rlmax)
    sysm.zero_forces()
    U, W, PT = sysm.force_sc(0, r_cut)
    F, T = sysm.get_forces(0)
    So = S.copy()
    lists = H.oracle_list(So, rlmax)
    rc, Uo, Wo, PTo, Fo, To = H.oracle_forces(So, "sc", lists=lists, r_cut=r_cut)
    assert rc == 0
    assert U == pytest.approx(Uo, rel=TOL64, abs=1e-13)
    assert W == pytest.approx(Wo, rel=TOL64, abs=1e-13)
    assert np.allclose(PT, PTo, rtol=TOL64, atol=1e-12 * (1 + np.abs(PTo).max()))
    if np.abs(Fo).max() > 0:
        assert H.max_rel_row(F, Fo) < TOL64
        assert H.max_rel_row(T, To) < TOL64
    else:
        assert np.abs(F).max() == 0


def test_force_sc_overlap_reported(msc):
    """Two rods 0.6 sigma apart: force_sc reports the overlap the reference
    aborts on (v25/functions.cpp:1454)."""
    from mdspherocyl_amd import replica_desc
    N = 2
    desc = replica_desc(N, 20.0, 0.3, 3, 4.5)
    sysm = _system(msc, desc)
    r = np.array([[0.0, 0.0, 0.0], [0.6, 0.0, 0.0]])
    ori = np.tile(np.eye(3), (N, 1, 1))
    sysm.set_state(0, r, np.zeros((N, 3)), np.zeros((N, 3)), ori, np.full(N, 3.0), np.ones(N), np.ones(N),
                   np.zeros((N, 1)), np.tile([0, 0, 1.0], (N, 1, 1)), np.ones((N, 1)))
    sysm.neigh_list(0, 6.9)
    sysm.zero_forces()
    with pytest.raises(msc.Overlap):
        sysm.force_sc(0, 4.5)


@pytest.mark.parametrize("case", ["v251", "c2"])
def test_force_dipR_fp64_all_pairs(msc, v251_init, c2_state, case):
    if case == "v251":
        S, P, rlmin, rlmax = v251_init
        desc, alpha = H.v251_desc(), 0.07
    else:
        desc, st, S = c2_state
        alpha = desc.alpha
    _say("case: force_dipR fp64 all pairs %s" % case)
    sysm = _system(msc, desc)
    H.load_osys(sysm, 0, S)
    sysm.zero_forces()
    U, PT = sysm.force_dipR(0)
    F, T = sysm.get_forces(0)
    Uo, PTo, Fo, To = H.oracle_forces(S.copy(), "R", alpha=alpha)
    assert U == pytest.approx(Uo, rel=TOL64)
    assert np.allclose(PT, PTo, rtol=TOL64, atol=TOL64 * np.abs(PTo).max())
    assert H.max_rel_row(F, Fo) < TOL64
    assert H.max_rel_row(T, To) < TOL64


@pytest.mark.parametrize("case", ["v251", "c2"])
def test_force_dipK_fp64(msc, v251_init, c2_state, case):
    if case == "v251":
        S, P, rlmin, rlmax = v251_init
        desc, alpha, K = H.v251_desc(), 0.07, 5
    else:
        desc, st, S = c2_state
        alpha, K = desc.alpha, desc.K
    _say("case: force_dipK fp64 %s K=%d" % (case, K))
    sysm = _system(msc, desc)
    H.load_osys(sysm, 0, S)
    sysm.zero_forces()
    U, PT = sysm.force_dipK(0)
    F, T = sysm.get_forces(0)
    Uo, PTo, Fo, To = H.oracle_forces(S.copy(), "K", alpha=alpha, K=K)
    assert U == pytest.approx(Uo, rel=TOL64)
    assert np.allclose(PT, PTo, rtol=1e-10, atol=1e-10 * np.abs(PTo).max())
    assert H.max_rel_row(F, Fo) < 1e-10
    assert H.max_rel_row(T, To) < 1e-10


def test_ext_field_and_self_energy(msc, v251_init):
    S, P, rlmin, rlmax = v251_init
    _say("case: ext_field / self_ener v251")
    sysm = _system(msc, H.v251_desc())
    H.load_osys(sysm, 0, S)
    sysm.zero_forces()
    U = sysm.ext_field(0)
    F, T = sysm.get_forces(0)
    Uo, _, Fo, To = H.oracle_forces(S.copy(), "ext", B=(0.0, 0.0, 5.0))
    assert U == pytest.approx(Uo, rel=1e-13)
    assert np.allclose(T, To, rtol=1e-13, atol=1e-14)
    assert sysm.self_ener(0) == 0.0


def test_cutoff_mode_fp64(msc, c2_state):
    """Real space restricted to dipole pairs within r_c = 3.5 (pair list built
    with the neighbour list) against the oracle's cut-off sum."""
    desc, st, S = c2_state
    from mdspherocyl_amd import replica_desc
    d2 = replica_desc(desc.N, desc.L[0], 3.0 / 3.5, 8, desc.sr_cut, rc=3.5)
    _say("case: force_dipR fp64 cut-off 3.5, c2 state")
    sysm = _system(msc, d2)
    H.load_osys(sysm, 0, S)
    rlmin, rlmax = H.list_radii(S, desc.sr_cut)
    sysm.neigh_list(0, rlmax)
    sysm.zero_forces()
    U, PT = sysm.force_dipR(0)
    F, T = sysm.get_forces(0)
    Uo, PTo, Fo, To = H.oracle_forces(S.copy(), "R", alpha=3.0 / 3.5, rc=3.5)
    assert U == pytest.approx(Uo, rel=TOL64)
    assert H.max_rel_row(F, Fo) < TOL64
    assert H.max_rel_row(T, To) < TOL64


@pytest.mark.parametrize("part", ["R", "K", "Rcut"])
def test_mixed_precision_errors(msc, c2_state, part):
    desc, st, S = c2_state
    from mdspherocyl_amd import replica_desc
    if part == "Rcut":
        d2 = replica_desc(desc.N, desc.L[0], 3.0 / 3.5, 8, desc.sr_cut, rc=3.5)
    else:
        d2 = desc
    _say("case: mixed precision %s, c2 state" % part)
    sysm = _system(msc, d2, "mixed")
    H.load_osys(sysm, 0, S)
    if part == "Rcut":
        rlmin, rlmax = H.list_radii(S, desc.sr_cut)
        sysm.neigh_list(0, rlmax)
    sysm.zero_forces()
    if part == "K":
        U, PT = sysm.force_dipK(0)
        Uo, PTo, Fo, To = H.oracle_forces(S.copy(), "K", alpha=d2.alpha, K=d2.K)
    else:
        U, PT = sysm.force_dipR(0)
        Uo, PTo, Fo, To = H.oracle_forces(S.copy(), "R", alpha=d2.alpha, rc=d2.rc)
    F, T = sysm.get_forces(0)
    eF, eT = H.rel_err(F, Fo), H.rel_err(T, To)
    _say("  rel L2 error F %.3e T %.3e U %.3e" % (eF, eT, abs(U - Uo) / abs(Uo)))
    assert eF <= 1e-5 and eT <= 1e-5
    assert U == pytest.approx(Uo, rel=1e-6)


def _fixture_lines():
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "v251_run1_output.txt")
    return [[float(x) for x in l.split("|")[1].split()] for l in open(path).read().splitlines()]


def test_step_loop_reproduces_reference_run(msc, v251_init):
    """101 steps of v25/v251/run1 on the GPU (FP64): every printed observable
    of the reference's output.dat within 1e-9 relative."""
    S, P, rlmin, rlmax = v251_init
    _say("case: step loop v251, 101 steps fp64")
    sysm = _system(msc, H.v251_desc(alpha=0.07, K=5))
    H.load_osys(sysm, 0, S)
    sysm.init_run()
    obs = sysm.run(101)
    ref = _fixture_lines()
    for k, row in enumerate(ref):
        mine = obs[10 * k, 0, :11]
        for j in range(11):
            assert mine[j] == pytest.approx(row[j], rel=1e-9, abs=2e-12), "step %d column %d" % (10 * k, j)


def test_batched_replicas_equal_single_runs(msc):
    """Three replicas of different size in one context give the same forces
    as each replica alone."""
    from mdspherocyl_amd import configs, replica_desc
    rng = np.random.default_rng(11)
    reps = [configs.make_replica(rng, n, configs.truncated_normal(rng, n, 3.0, 0.3), 0.1, 1.5, 6,
                                 lambda b: 5.6 / b, rc) for n, rc in [(300, 0.0), (517, 3.0), (260, 0.0)]]
    _say("case: batched replicas N=300,517,260")
    big = msc.MdSystem([d for d, _ in reps])
    configs.load_into(big, reps)
    big.init_run()
    for k, (d, st) in enumerate(reps):
        one = msc.MdSystem([d])
        configs.load_into(one, [(d, st)])
        one.init_run()
        Fa, Ta = big.get_forces(k)
        Fb, Tb = one.get_forces(0)
        assert np.array_equal(Fa, Fb) and np.array_equal(Ta, Tb), "replica %d" % k
    obs = big.run(20)
    assert np.all(np.isfinite(obs))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nve_energy_drift(msc, precision):
    """c1-like system with the thermostat switched off (Q = 1e12, gamma = 0,
    as the reference's v25/run1): per-particle drift of K+U over 2000 steps
    stays below the reference's own 4e-3."""
    from mdspherocyl_amd import configs
    d, st = configs.build("c1")[0]
    d.Q = 1e12
    d.gamma = 0.0
    _say("case: NVE drift c1 %s" % precision)
    sysm = msc.MdSystem([d], precision=precision)
    configs.load_into(sysm, [(d, st)])
    sysm.init_run()
    obs = sysm.run(2000)
    E = obs[:, 0, 0] + obs[:, 0, 1]
    drift = float(np.max(np.abs(E - E[0])))
    _say("  max |E - E0| per particle = %.3e" % drift)
    assert drift < 4e-3
