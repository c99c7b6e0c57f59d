"""Shared builders for the parity tests: states made by the oracle (test
infrastructure) and handed identically to the HIP library."""
import ctypes
import math

import numpy as np

from oracle import oracle as O


def v251_state():
    """The v25/v251/run1 system after its Signal=0 initialisation
    (ini_pos, 1000 GCAmix, set_dip, ini_mom), as an oracle system."""
    N = 500
    S = O.OSys(N, 1)
    S.L[:] = math.pow(500.0 / ((1.0 * 1.0) * 0.0013), 1.0 / 3)
    S.len[:] = 9.0
    S.dia[:] = 1.0
    S.mass[:] = 1.0
    S.doc[:, :, :] = (0.0, 0.0, 1.0)
    S.dmu[:] = 4.481
    S.set_inertia()
    S.set_dip()
    P = O.run_params(dens=0.0013, temp=1.0, dt=1e-4, sr_cut=10.6, alpha=0.07, min_s=0.1, Q=5.0, gamma=1.0,
                     s0=1.0, B=(0.0, 0.0, 5.0), K=5, cl_fr=10)
    return S, P


def v25_state():
    """The system of v25/input.txt (two species of 50 rods, two off-centre
    tilted dipoles each, K = 10, alpha = 0.75) as an oracle system before its
    Signal=0 initialisation, its run parameters and the parsed input."""
    import os
    from mdspherocyl_amd import refio
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    inp = refio.read_input(os.path.join(gold, "v25_input.txt"))
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
    return S, P, inp


def v25_desc(S, inp, **over):
    """Library descriptor of the v25/input.txt run (real space over every
    minimum-image pair, as the reference sums)."""
    from mdspherocyl_amd import replica_desc
    kw = dict(rc=0.0, nd=S.nd, B=tuple(inp["B"]), temp=inp["temp"], dt=inp["Dt"], Q=inp["Q"], gamma=inp["gamma"],
              s0=0.0, dens=inp["dens"], thermo_seed=(1, 2))
    kw.update(over)
    return replica_desc(S.N, tuple(S.L), inp["alpha"], inp["K_max"], inp["sr_cut"], **kw)


def v251_desc(**over):
    from mdspherocyl_amd import replica_desc
    kw = dict(rc=0.0, B=(0.0, 0.0, 5.0), temp=1.0, dt=1e-4, Q=5.0, gamma=1.0, s0=1.0, dens=0.0013,
              thermo_seed=(1, 2))
    kw.update(over)
    L = math.pow(500.0 / ((1.0 * 1.0) * 0.0013), 1.0 / 3)
    return replica_desc(500, L, kw.pop("alpha", 0.07), kw.pop("K", 5), kw.pop("sr_cut", 10.6), **kw)


def osys_from_config(desc, st):
    """Oracle copy of a configs.build replica."""
    N = desc.N
    S = O.OSys(N, 1, L=tuple(desc.L))
    S.r[:] = st["r"]
    S.p[:] = st["p"]
    S.l[:] = st["l"]
    S.ori[:] = st["ori"]
    S.len[:] = st["length"]
    S.dia[:] = st["dia"]
    S.mass[:] = st["mass"]
    S.dpc[:] = st["dpc"]
    S.doc[:] = st["doc"]
    S.dmu[:] = st["dmu"]
    S.set_inertia()
    S.set_dip()
    return S


def load_osys(system, rep, S):
    system.set_state(rep, S.r, S.p, S.l, S.ori, S.len, S.dia, S.mass, S.dpc, S.doc, S.dmu)


def list_radii(S, sr_cut):
    """r_list_min / r_list_max of v25/main.cpp:461-463 for per-particle shapes."""
    dmax = float((S.len + S.dia).max())
    dmin = float(S.dia.min())
    rmin = sr_cut + 0.1 * dmin
    if rmin < 1.5 * dmax:
        rmin = 1.5 * dmax
    return rmin, rmin + 0.9 * dmin


def oracle_list(S, r_list):
    L = O.lib()
    c = S.c()
    tot = L.o_neigh_list(c, r_list, None, None, 0)
    off = np.zeros(S.N + 1, dtype=np.int32)
    nbr = np.zeros(max(tot, 1), dtype=np.int32)
    L.o_neigh_list(S.c(), r_list, O.iptr(off), O.iptr(nbr), tot)
    return off, nbr[:tot]


def oracle_forces(S, which, lists=None, alpha=None, K=None, rc=0.0, r_cut=None, B=(0.0, 0.0, 0.0)):
    L = O.lib()
    S.F[:] = 0
    S.T[:] = 0
    U, W, PT = ctypes.c_double(), ctypes.c_double(), np.zeros(9)
    if which == "sc":
        off, nbr = lists
        rc_ = L.o_force_sc(S.c(), r_cut, O.iptr(off), O.iptr(nbr), ctypes.byref(U), ctypes.byref(W), O.dptr(PT))
        return rc_, U.value, W.value, PT.reshape(3, 3), S.F.copy(), S.T.copy()
    if which == "R":
        T = np.zeros(32768 * 3)
        L.o_tabul(O.dptr(T), alpha)
        L.o_force_dipR(S.c(), O.dptr(T), rc, ctypes.byref(U), O.dptr(PT))
    elif which == "K":
        KV = np.zeros((K + 1) ** 3)
        L.o_wave_v(O.dptr(KV), O.dptr(S.L), alpha, K)
        L.o_force_dipK(S.c(), K, alpha, O.dptr(KV), ctypes.byref(U), O.dptr(PT))
    elif which == "ext":
        b = np.array(B, dtype=np.float64)
        L.o_ext_field(S.c(), O.dptr(b), ctypes.byref(U))
    return U.value, PT.reshape(3, 3), S.F.copy(), S.T.copy()


def rel_err(a, b):
    """||a - b|| / ||b|| over all components."""
    a, b = np.asarray(a), np.asarray(b)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def max_rel_row(a, b):
    """max_i |a_i - b_i| / rms_i |b_i| (per-particle vector error against the
    typical magnitude)."""
    a, b = np.asarray(a), np.asarray(b)
    rms = np.sqrt(np.mean(np.sum(b * b, axis=1)))
    err = float(np.max(np.linalg.norm(a - b, axis=1)))
    return err / rms if rms > 0 else err
