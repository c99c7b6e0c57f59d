"""Pin the oracle against outputs the reference itself wrote.

The reference run v25/v251/run1 (N=500 spherocylinders, L=9, one axial dipole
mu=4.481, B=(0,0,5), K=5, alpha=0.07, NHL thermostat gamma=1, dt=1e-4, 101
steps, clustering every 10 steps) is replayed from its Signal=0 start:
ini_pos, 1000 GCAmix moves, set_dip, ini_mom, then the step loop of
v25/main.cpp.  Every printed value is compared at the printed precision
(12 decimals): observables (output.dat), the final state and cluster labels
(conf.dat), 25-particle snapshots every 2 steps (pos_vel_*.dat), the
neighbour-list reference positions (compConf.dat) and the lattice (iniPos.dat).
"""
import ctypes
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
N = 500


def v251_system():
    S = O.OSys(N, 1)
    S.L[:] = math.pow(500.0 / ((1.0 * 1.0) * 0.0013), 1.0 / 3)
    S.len[:] = 9.0
    S.dia[:] = 1.0
    S.mass[:] = 1.0
    S.doc[:, :, :] = (0.0, 0.0, 1.0)
    S.dmu[:] = 4.481
    S.set_inertia()
    S.set_dip()
    P = O.run_params(dens=0.0013, temp=1.0, dt=1e-4, sr_cut=10.6, alpha=0.07, min_s=0.1,
                     Q=5.0, gamma=1.0, s0=1.0, B=(0.0, 0.0, 5.0), K=5, cl_fr=10)
    return S, P


def f12(v):
    return "%.12f" % v


def mm(A, B):
    return [[sum_ordered(A[m][i] * B[i][n] for i in range(len(B))) for n in range(len(B[0]))]
            for m in range(len(A))]


def sum_ordered(it):
    acc = 0.0
    for x in it:
        acc += x
    return acc


def posvel_line(S, i):
    """pos_vel line of particle i (v25/functions.cpp:2602-2643)."""
    m = S.mass[i]
    r, p = S.r[i], S.p[i]
    o = [list(row) for row in S.ori[i]]
    I = S.inert[i]
    I0 = [[1.0 / I[0], 0.0, 0.0], [0.0, 1.0 / I[1], 0.0], [0.0, 0.0, 1.0 / I[2]]]
    ot = [[o[n][m_] for n in range(3)] for m_ in range(3)]
    Ib = mm(mm(ot, I0), o)
    om = mm(Ib, [[S.l[i][0]], [S.l[i][1]], [S.l[i][2]]])
    vals = [r[0], r[1], r[2], p[0] / m, p[1] / m, p[2] / m]
    vals2 = [o[2][0], o[2][1], o[2][2], om[0][0], om[1][0], om[2][0]]
    return " ".join(f12(v) for v in vals) + " | " + " ".join(f12(v) for v in vals2)


def helfand(S):
    """print_ptens's Helfand sum (v25/v251/functions.cpp:2646-2665):
    sum over particles of r_a p_b, summed in particle order."""
    h = [[0.0] * 3 for _ in range(3)]
    for i in range(S.N):
        r, p = S.r[i].tolist(), S.p[i].tolist()
        for a in range(3):
            for b in range(3):
                h[a][b] += r[a] * p[b]
    return [v for row in h for v in row]


@pytest.fixture(scope="module")
def v251_run():
    S, P = v251_system()
    R = O.ORun(S, P, seed_init=True)
    posvel = {}
    obs = np.zeros((101, O.OBS_N))
    nb = int(math.floor(math.sqrt(3.) * S.L.min() / (2. * 0.01)))
    hist = np.zeros((1, 4, nb))
    dropped = 0
    oo, pt = [], []
    for step in range(101):
        obs[step] = R.run(1)[0]
        if step % 2 == 0:
            posvel[step // 2] = [posvel_line(S, i) for i in range(25)]
        if step % 10 == 0:  # corrf = 10, after the step's clustering (v25/main.cpp:710-716)
            dropped += O.lib().o_acc_correl(S.c(), 0.01, nb, O.dptr(hist))
        g1, g2, M = ctypes.c_double(), ctypes.c_double(), np.zeros(3)
        O.lib().o_oo_params(S.c(), ctypes.byref(g1), ctypes.byref(g2), O.dptr(M))
        oo.append((g1.value, g2.value, M))
        pt.append((helfand(S), list(R.R.PT)))
    S.correl = (hist, dropped)
    S.oo = oo
    S.pt = pt
    yield S, P, R, obs, posvel
    R.free()


def test_output_dat_observables(v251_run):
    S, P, R, obs, _ = v251_run
    lines = open(os.path.join(GOLD, "v251_run1_output.txt")).read().splitlines()
    assert len(lines) == 11
    for k, line in enumerate(lines):
        ref = line.split("|")[1].split()
        mine = [f12(v) for v in obs[10 * k, :11]]
        assert mine == ref, "step %d" % (10 * k)


def test_conf_dat_state_and_cluster_labels(v251_run):
    S, P, R, obs, _ = v251_run
    lines = open(os.path.join(GOLD, "v251_run1_conf.txt")).read().splitlines()
    nlab = 0
    for i in range(N):
        tok = lines[1 + i].split()
        bar = tok.index("|")
        mine = list(S.r[i]) + list(S.p[i]) + list(S.ori[i].ravel()) + list(S.l[i])
        assert [f12(v) for v in mine] == tok[:bar], "particle %d" % i
        assert int(tok[-1]) == S.clust[i], "cluster label of particle %d" % i
        nlab = max(nlab, int(tok[-1]) + 1)
    assert nlab == 478
    for k in range(2):
        s, strn = lines[501 + k].split()
        assert f12(R.R.th[k].s) == s and f12(R.R.th[k].strn) == strn


def test_pos_vel_snapshots(v251_run):
    S, P, R, obs, posvel = v251_run
    blocks = {}
    cur = None
    for line in open(os.path.join(GOLD, "v251_run1_posvel.txt")).read().splitlines():
        if line.startswith("# file"):
            cur = int(line.split()[2])
            blocks[cur] = []
        else:
            blocks[cur].append(line)
    assert len(blocks) == 51
    for n in range(51):
        assert posvel[n] == blocks[n][:25], "pos_vel_%d" % n
        assert blocks[n][25] == f12(2 * 1e-4)


def test_compconf_neighbour_reference(v251_run):
    S, P, R, obs, _ = v251_run
    ref = open(os.path.join(GOLD, "v251_run1_compConf.txt")).read().splitlines()
    snap = np.ctypeslib.as_array(R.R.snap, shape=(N * 3,)).reshape(N, 3)
    assert [" ".join(f12(v) for v in snap[i]) for i in range(N)] == ref


def test_ini_pos_lattice():
    S, P = v251_system()
    assert O.lib().o_ini_pos(S.c()) == 0
    ref = open(os.path.join(GOLD, "v251_run1_iniPos.txt")).read().splitlines()
    for i in range(N):
        tok = ref[i].split()
        assert ["%g" % v for v in S.r[i]] == tok[:3], "particle %d" % i
    assert ref[N].split() == ["%g" % v for v in S.L]


def test_nve_drift_statistics_of_reference_run():
    """v25/run1 (N=100 dipolar spheres, Q=1e12, gamma=0) conserves K+U per
    particle within 4e-3 over 20001 steps; the figure the GPU NVE test is
    held to is read from here."""
    d = np.loadtxt(os.path.join(GOLD, "v25_run1_nve.txt"))
    seg1 = d[d[:, 0] <= 6089]
    E = seg1[:, 2] + seg1[:, 3]
    drift = np.max(np.abs(E - E[0]))
    assert 2e-3 < drift < 4e-3


def test_correlation_dat(v251_run):
    """acc_correl at steps 0, 10, ..., 100 of the replayed run, normalised
    as nor_correl does: every line of the reference's correlation.dat
    (6298 bars, 8 decimals), the first columns byte for byte and each
    histogram value within one unit of its last printed digit (the
    reference sums with OpenMP atomics in no fixed order)."""
    import gzip
    from mdspherocyl_amd import refio
    S = v251_run[0]
    hist, dropped = S.correl
    assert dropped == 0
    V = S.L[0] * S.L[1] * S.L[2]
    cor = 0.0
    for _ in range(11):
        cor += 1. / V
    mine = refio.correlation_text(hist, [500], S.L.min(), 0.01, cor).splitlines()
    ref = gzip.open(os.path.join(GOLD, "v251_run1_correlation.txt.gz"), "rt").read().splitlines()
    assert len(mine) == len(ref) == 6298
    same = 0
    for a, b in zip(mine, ref):
        ha, ta = a.split("&")
        hb, tb = b.split("&")
        assert ha == hb
        va = [float(x) for x in ta.replace("|", " ").split()]
        vb = [float(x) for x in tb.replace("|", " ").split()]
        assert np.allclose(va, vb, rtol=0, atol=1.01e-8), a + "\n" + b
        same += a == b
    assert same >= 6290
    # the counts channel: exact pair counts, 2 per pair and frame
    assert hist[0, 0].sum() == 2 * 11 * (500 * 499 // 2)


def test_magmom_dat(v251_run):
    """oo_params' total dipole moment after each of the 101 steps equals
    magmom.dat (tcfv = 1) to its 12 decimals.  M_z is about 2240, where the
    12th decimal is two units of the last place of a double: there a value
    may sit one printed unit off (one step of the 101 does)."""
    S = v251_run[0]
    lines = open(os.path.join(GOLD, "v251_run1_magmom.txt")).read().splitlines()
    assert len(lines) == 101 == len(S.oo)
    exact = 0
    for k, line in enumerate(lines):
        ref = line.split()[1:]
        mine = ["%.12f" % v for v in S.oo[k][2]]
        exact += mine == ref
        assert np.all(np.abs(np.array(S.oo[k][2]) - np.array([float(x) for x in ref])) <= 1.01e-12), k
    assert exact >= 100


def test_counters_order_parameters(v251_run):
    """The sums of G1, G1^2, G2 and G2^2 over the 101 steps (single-precision
    Q tensor and Jacobi rotations as the reference's build runs them) equal
    the four order-parameter fields of counters.dat to 12 decimals."""
    S = v251_run[0]
    tok = open(os.path.join(GOLD, "v251_run1_counters.txt")).read().split()
    sums = [0.0] * 4
    for g1, g2, _ in S.oo:
        sums[0] += g1
        sums[1] += g1 * g1
        sums[2] += g2
        sums[3] += g2 * g2
    assert ["%.12f" % v for v in sums] == tok[115:119]


def test_jacobi3_closed_forms():
    """The Jacobi routine on matrices with known spectra: a rotated diagonal
    matrix gives its eigenvalues and (up to sign) the rotation's columns; a
    diagonal one needs no rotation; the eigenvectors are orthonormal."""
    lib = O.lib()
    fp = ctypes.POINTER(ctypes.c_float)
    rng = np.random.default_rng(5)
    for trial in range(20):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        lam = np.array([0.9, -0.35, -0.55]) if trial % 2 else rng.normal(size=3)
        A = (q * lam) @ q.T
        a = np.ascontiguousarray(A, dtype=np.float32)
        d, v = np.zeros(3, np.float32), np.zeros((3, 3), np.float32)
        nrot = lib.o_jacobi3(a.ctypes.data_as(fp), d.ctypes.data_as(fp), v.ctypes.data_as(fp))
        assert nrot > 0
        assert np.allclose(np.sort(d), np.sort(lam), atol=2e-6)
        assert np.allclose(v.T @ v, np.eye(3), atol=2e-6)
        for k in range(3):
            assert np.allclose(A.astype(np.float32) @ v[:, k], d[k] * v[:, k], atol=5e-6)
    a = np.diag([0.5, -0.25, -0.25]).astype(np.float32)
    d, v = np.zeros(3, np.float32), np.zeros((3, 3), np.float32)
    assert lib.o_jacobi3(a.ctypes.data_as(fp), d.ctypes.data_as(fp), v.ctypes.data_as(fp)) == 0
    assert list(d) == [0.5, -0.25, -0.25] and np.array_equal(v, np.eye(3, dtype=np.float32))


def test_ptens_dat(v251_run):
    """ptens.dat (tcfv = 1: every step): the Helfand sum r_a p_b and the total
    pressure tensor (PTkin + PTss + PTdipR + PTdipK) / V of
    v25/v251/main.cpp:635-642, including the reciprocal part's
    a4 = 1/(4 alpha) (v25/functions.cpp:1715, 1738), equal the file's 18
    values at its 12 decimals.  The Helfand sums reach 764, where the 12th
    decimal is eight units of the last place of a double: a value may sit one
    printed unit off there."""
    S = v251_run[0]
    lines = open(os.path.join(GOLD, "v251_run1_ptens.txt")).read().splitlines()
    assert len(lines) == 101 == len(S.pt)
    exact = 0
    for k, line in enumerate(lines):
        ref = line.split()[1:]
        assert len(ref) == 18
        hf, pt = S.pt[k]
        mine = [f12(v) for v in hf + pt]
        exact += mine == ref
        # the tensor part exactly; the Helfand part within one printed unit
        assert mine[9:] == ref[9:], "pressure tensor of step %d" % k
        assert np.all(np.abs(np.array(hf) - np.array([float(x) for x in ref[:9]])) <= 1.01e-12), k
    assert exact >= 99


def test_v25_two_species_two_dipole_trajectory():
    """v25/input.txt from its Signal=0 start: two species of 50 rods (L = D =
    1), each with two off-centre dipoles (at -0.5 and +0.5 along the axis,
    tilted in the body frame, moments 1.5 and 2), K = 10, alpha = 0.75,
    NHL thermostat.  All 661 steps of the run printed at v25/output.dat
    lines 1220-1880 (K U P Uss Pss UdipR PdipR UdipK PdipK |P|^2 |L|^2)
    equal the oracle's at 12 decimals.  This pins the lever-arm terms of the
    real-space and reciprocal sums, the reciprocal intra-particle exclusion
    over two dipole slots, the per-species inertia and ini_pos / GCAmix /
    ini_mom for two species (including ini_mom's reads past its per-species
    arrays, v25/functions.cpp:717-726)."""
    from mdspherocyl_amd import refio
    inp = refio.read_input(os.path.join(GOLD, "v25_input.txt"))
    assert inp["signal"] == 0 and inp["n_part"] == [50, 50]
    N = sum(inp["n_part"])
    Lb = math.pow(N / (inp["ArLx"] * inp["ArLy"] * inp["dens"]), 1.0 / 3)
    nd = len(inp["dip"][0])
    S = O.OSys(N, nd, L=(inp["ArLx"] * Lb, inp["ArLy"] * Lb, Lb), counts=inp["n_part"])
    b = 0
    for n, NN in enumerate(inp["n_part"]):
        S.len[b:b + NN] = inp["length"][n]
        S.dia[b:b + NN] = inp["diameter"][n]
        S.mass[b:b + NN] = inp["mass"][n]
        for w, (pos, ori, mu) in enumerate(inp["dip"][n]):
            S.dpc[b:b + NN, w] = pos
            S.doc[b:b + NN, w] = ori
            S.dmu[b:b + NN, w] = mu
        b += NN
    S.set_inertia()
    P = O.run_params(dens=inp["dens"], temp=inp["temp"], dt=inp["Dt"], sr_cut=inp["sr_cut"], alpha=inp["alpha"],
                     min_s=inp["min_s"], Q=inp["Q"], gamma=inp["gamma"], s0=0.0, B=tuple(inp["B"]),
                     K=inp["K_max"], cl_fr=inp["cl_fr"])
    R = O.ORun(S, P, seed_init=True)
    try:
        obs = R.run(661)
    finally:
        R.free()
    lines = open(os.path.join(GOLD, "v25_output_seg.txt")).read().splitlines()
    assert len(lines) == 661
    for k, line in enumerate(lines):
        head, tail = line.split("|")
        # the time column restarts after the pre_steps equilibration steps
        t = (k + 1 if k < inp["pre_steps"] else k + 1 - inp["pre_steps"]) * inp["Dt"]
        assert abs(float(head.split()[0].rstrip(".")) - t) < 1e-12
        assert [f12(v) for v in obs[k, :11]] == tail.split(), "step %d" % (k + 1)
