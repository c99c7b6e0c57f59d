"""Host-side mirror of the reference's force-path routines over the C ABI.

``MdSystem`` owns one library context (one HIP stream on one GPU) holding one
or more independent replicas.  Its methods carry the names and argument
meaning of the reference routines they replace (v25/functions.hpp:29-54):
force routines add into F and T and return energies as totals; ``force_sc``
raises ``Overlap`` where the reference returns false.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (MSC_ALL, MSC_E_DISTANCE, MSC_E_OVERLAP, MSC_FP64, MSC_MIXED, OBS_N, MscError,
                   ReplicaDesc)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Overlap(MscError):
    """force_sc found d^2 < 0.49 sigma^2 or d^2 > 1.00001 r^2 (v25/functions.cpp:1454-1473)."""


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def replica_desc(N, L, alpha, K, sr_cut, rc=0.0, nd=1, B=(0.0, 0.0, 0.0), temp=1.0, dt=0.002, Q=5.0,
                 gamma=0.0, s0=0.0, dens=0.0, dmin=0.0, dmax=0.0, thermo_seed=(1, 2)):
    d = ReplicaDesc()
    d.N, d.nd = int(N), int(nd)
    L = np.broadcast_to(np.asarray(L, dtype=np.float64), (3,))
    for a in range(3):
        d.L[a] = L[a]
        d.B[a] = B[a]
    d.alpha, d.K, d.sr_cut, d.rc = alpha, int(K), sr_cut, rc
    d.temp, d.dt, d.Q, d.gamma, d.s0, d.dens = temp, dt, Q, gamma, s0, dens
    d.dmin, d.dmax = dmin, dmax
    d.thermo_seed[0], d.thermo_seed[1] = int(thermo_seed[0]), int(thermo_seed[1])
    return d


class MdSystem:
    """Replicas of mdSpherocyl systems advanced on one MI355X."""

    def __init__(self, descs, precision="fp64", device=0):
        self.L = _lib.load()
        if isinstance(descs, ReplicaDesc):
            descs = [descs]
        self.descs = list(descs)
        self.nrep = len(self.descs)
        self.precision = MSC_MIXED if precision in ("mixed", MSC_MIXED) else MSC_FP64
        arr = (ReplicaDesc * self.nrep)(*self.descs)
        self._ctx = ctypes.c_void_p()
        rc = self.L.msc_create(device, self.precision, self.nrep, arr, ctypes.byref(self._ctx))
        if rc:
            raise MscError(rc, "msc_create failed: " + (self.L.msc_last_error(None) or b"").decode())

    # -- plumbing ----------------------------------------------------------
    def _chk(self, rc):
        if rc:
            msg = self.L.msc_last_error(self._ctx).decode()
            if rc in (MSC_E_OVERLAP, MSC_E_DISTANCE):
                raise Overlap(rc, msg)
            raise MscError(rc, msg)

    def close(self):
        if self._ctx:
            self.L.msc_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state ---------------------------------------------------------------
    def set_state(self, rep, r, p, l, ori, length, dia, mass, dpc, doc, dmu):
        N = self.descs[rep].N
        nd = self.descs[rep].nd
        arrs = [np.reshape(r, (N, 3)), np.reshape(p, (N, 3)), np.reshape(l, (N, 3)), np.reshape(ori, (N, 3, 3)),
                np.reshape(length, (N,)), np.reshape(dia, (N,)), np.reshape(mass, (N,)),
                np.reshape(dpc, (N, nd)), np.reshape(doc, (N, nd, 3)), np.reshape(dmu, (N, nd))]
        keep = [_d(a) for a in arrs]
        self._chk(self.L.msc_set_state(self._ctx, rep, *[k[1] for k in keep]))

    def get_state(self, rep):
        N = self.descs[rep].N
        r, p, l, o = np.zeros((N, 3)), np.zeros((N, 3)), np.zeros((N, 3)), np.zeros((N, 3, 3))
        self._chk(self.L.msc_get_state(self._ctx, rep, *[_d(a)[1] for a in (r, p, l, o)]))
        return r, p, l, o

    def get_forces(self, rep):
        N = self.descs[rep].N
        F, T = np.zeros((N, 3)), np.zeros((N, 3))
        self._chk(self.L.msc_get_forces(self._ctx, rep, _d(F)[1], _d(T)[1]))
        return F, T

    def set_forces(self, rep, F, T):
        self._chk(self.L.msc_set_forces(self._ctx, rep, _d(F)[1], _d(T)[1]))

    def thermostat(self, rep):
        s, strn = np.zeros(2), np.zeros(2)
        self._chk(self.L.msc_get_thermostat(self._ctx, rep, _d(s)[1], _d(strn)[1]))
        return s, strn

    def set_thermostat(self, rep, s, strn):
        a, b = _d(s), _d(strn)
        self._chk(self.L.msc_set_thermostat(self._ctx, rep, a[1], b[1]))

    # -- reference routines ------------------------------------------------
    def neigh_list(self, rep, r_list):
        self._chk(self.L.msc_neigh_list(self._ctx, rep, float(r_list)))

    def nblist(self, rep):
        """The list in the reference's form (CSR of ascending g > f)."""
        N = self.descs[rep].N
        off = np.zeros(N + 1, dtype=np.int32)
        tot = ctypes.c_long()
        self._chk(self.L.msc_get_neigh_list(self._ctx, rep, off.ctypes.data_as(_ip), None, 0, ctypes.byref(tot)))
        nbr = np.zeros(max(tot.value, 1), dtype=np.int32)
        self._chk(self.L.msc_get_neigh_list(self._ctx, rep, off.ctypes.data_as(_ip), nbr.ctypes.data_as(_ip),
                                            tot.value, ctypes.byref(tot)))
        return off, nbr[:tot.value]

    def set_nblist(self, rep, r_list, off, nbr):
        """Make the caller's list (CSR of ascending g > f, the reference's
        nblist) the replica's neighbour rows."""
        off = np.ascontiguousarray(off, dtype=np.int32)
        nbr = np.ascontiguousarray(nbr, dtype=np.int32)
        if nbr.size == 0:
            nbr = np.zeros(1, dtype=np.int32)
        self._chk(self.L.msc_set_neigh_list(self._ctx, rep, float(r_list), off.ctypes.data_as(_ip),
                                            nbr.ctypes.data_as(_ip)))

    def check_list(self, rep, r_list_min, r_list_max):
        out = ctypes.c_int()
        self._chk(self.L.msc_check_list(self._ctx, rep, r_list_min, r_list_max, ctypes.byref(out)))
        return bool(out.value)

    def zero_forces(self, rep=MSC_ALL):
        self._chk(self.L.msc_zero_forces(self._ctx, rep))

    def force_sc(self, rep, r_cut):
        U, W, PT = ctypes.c_double(), ctypes.c_double(), np.zeros(9)
        self._chk(self.L.msc_force_sc(self._ctx, rep, r_cut, ctypes.byref(U), ctypes.byref(W), _d(PT)[1]))
        return U.value, W.value, PT.reshape(3, 3)

    def force_dipR(self, rep):
        U, PT = ctypes.c_double(), np.zeros(9)
        self._chk(self.L.msc_force_dipR(self._ctx, rep, ctypes.byref(U), _d(PT)[1]))
        return U.value, PT.reshape(3, 3)

    def force_dipK(self, rep):
        U, PT = ctypes.c_double(), np.zeros(9)
        self._chk(self.L.msc_force_dipK(self._ctx, rep, ctypes.byref(U), _d(PT)[1]))
        return U.value, PT.reshape(3, 3)

    def ext_field(self, rep):
        U = ctypes.c_double()
        self._chk(self.L.msc_ext_field(self._ctx, rep, ctypes.byref(U)))
        return U.value

    def self_ener(self, rep):
        U = ctypes.c_double()
        self._chk(self.L.msc_self_ener(self._ctx, rep, ctypes.byref(U)))
        return U.value

    def moveX(self, rep, dt):
        self._chk(self.L.msc_moveX(self._ctx, rep, dt))

    def moveP(self, rep, dt, last):
        K, PT = np.zeros(2), np.zeros(9)
        self._chk(self.L.msc_moveP(self._ctx, rep, dt, int(bool(last)), _d(K)[1], _d(PT)[1]))
        return K, PT.reshape(3, 3)

    def renorm(self, rep=MSC_ALL):
        self._chk(self.L.msc_renorm(self._ctx, rep))

    # -- step loop -----------------------------------------------------------
    def init_run(self):
        self._chk(self.L.msc_init_run(self._ctx))

    def run(self, nsteps, step0=0, observe=True):
        obs = np.zeros((max(nsteps, 1), self.nrep, OBS_N)) if observe else None
        self._chk(self.L.msc_run(self._ctx, step0, nsteps, _d(obs)[1] if observe else None))
        return obs[:nsteps] if observe else None

    def set_contact_rows(self, on):
        """WCA over the contact rows (True; mixed contexts' default) or over
        every reference-row pair (False; FP64 contexts' default)."""
        self._chk(self.L.msc_set_contact_rows(self._ctx, 1 if on else 0))

    def set_graph(self, on):
        self._chk(self.L.msc_set_graph(self._ctx, int(bool(on))))

    def get_tuning(self):
        """Launch choices (msc_tuning) as a dict."""
        t = _lib.Tuning()
        self._chk(self.L.msc_get_tuning(self._ctx, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, graph=None, fused_pairs=None, pair_waves=None, list_skin=None, chunk=None, overlap=None,
                   fuse_kick=None):
        t = _lib.Tuning()
        self._chk(self.L.msc_get_tuning(self._ctx, ctypes.byref(t)))
        if graph is not None:
            t.graph = int(bool(graph))
        if fused_pairs is not None:
            t.fused_pairs = int(bool(fused_pairs))
        if pair_waves is not None:
            t.pair_waves = int(pair_waves)
        if list_skin is not None:
            t.list_skin = float(list_skin)
        if chunk is not None:
            t.chunk = int(chunk)
        if overlap is not None:
            t.overlap = int(bool(overlap))
        if fuse_kick is not None:
            t.fuse_kick = int(bool(fuse_kick))
        self._chk(self.L.msc_set_tuning(self._ctx, ctypes.byref(t)))

    def autotune(self, steps=100):
        """msc_autotune: times every candidate launch choice over `steps`
        steps from the current state (restored after each trial) and keeps
        the fastest.  Returns (choice dict, [(candidate dict, ms/step)])."""
        best = _lib.Tuning()
        tried = (_lib.Tuning * 32)()
        ms = (ctypes.c_double * 32)()
        n = ctypes.c_int()
        self._chk(self.L.msc_autotune(self._ctx, int(steps), ctypes.byref(best), tried, ms, 32, ctypes.byref(n)))
        return best.as_dict(), [(tried[i].as_dict(), ms[i]) for i in range(min(n.value, 32))]

    # -- measurement --------------------------------------------------------
    def set_timing(self, on):
        self._chk(self.L.msc_set_timing(self._ctx, int(bool(on))))

    def reset_timing(self):
        self._chk(self.L.msc_reset_timing(self._ctx))

    def kernel_time(self, name):
        ms, n = ctypes.c_double(), ctypes.c_long()
        self._chk(self.L.msc_kernel_time(self._ctx, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def chunk_repeats(self):
        n = ctypes.c_long()
        self._chk(self.L.msc_chunk_repeats(self._ctx, ctypes.byref(n)))
        return n.value

    def order_params(self, rep):
        """oo_params (v25/functions.cpp:2702-2760): (G1, G2, M) of the
        replica, M the total dipole moment."""
        g1, g2, M = ctypes.c_double(), ctypes.c_double(), np.zeros(3)
        self._chk(self.L.msc_order_params(self._ctx, rep, ctypes.byref(g1), ctypes.byref(g2), _d(M)[1]))
        return g1.value, g2.value, M

    def order_params_ref(self, rep):
        """oo_params as the reference computes it (single-precision Q tensor
        and jacobi, sums in particle order): the reference's (G1, G2, M) bit
        for bit given the same state, for output-file parity."""
        g1, g2, M = ctypes.c_double(), ctypes.c_double(), np.zeros(3)
        self._chk(self.L.msc_order_params_ref(self._ctx, rep, ctypes.byref(g1), ctypes.byref(g2), _d(M)[1]))
        return g1.value, g2.value, M

    def pair_correl(self, rep, bar_w, n_bars, counts=None, hist=None):
        """acc_correl (v25/functions.cpp:2399-2466): adds this replica's pair
        histograms into ``hist`` ([sets][4][n_bars], new zeros when None)
        and returns (hist, dropped); ``counts`` are the species sizes (one
        species when None)."""
        N = self.descs[rep].N
        counts = np.ascontiguousarray([N] if counts is None else counts, dtype=np.int32)
        nspec = len(counts)
        sets = nspec + nspec * (nspec - 1) // 2
        if hist is None:
            hist = np.zeros((sets, 4, n_bars))
        if hist.dtype != np.float64 or not hist.flags["C_CONTIGUOUS"] or hist.size != sets * 4 * n_bars:
            raise ValueError("hist must be a C-contiguous float64 array of %d values" % (sets * 4 * n_bars))
        dropped = ctypes.c_long(0)
        self._chk(self.L.msc_pair_correl(self._ctx, rep, nspec, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                         bar_w, n_bars, _d(hist)[1], ctypes.byref(dropped)))
        return hist, dropped.value

    def cluster_labels(self, rep, min_s, mt_state, mt_index, u_cut=None):
        """clust_conf (v25/functions.cpp:3270): per-particle cluster numbers.

        mt_state (624 words) and mt_index are the MT19937 state of the
        reference's `ran` generator; returns (labels, nclust, mt_state,
        mt_index, n_asym) with the state advanced by the draws made.
        ``u_cut`` (None: the reference's geometric test) adds the dipolar
        clause: a contact also needs a bare dipolar pair energy below u_cut.
        """
        st = np.ascontiguousarray(mt_state, dtype=np.uint32).copy()
        if st.shape != (624,):
            raise ValueError("mt_state must hold 624 words")
        idx = ctypes.c_int(int(mt_index))
        n = self.descs[rep].N
        labels = np.zeros(n, dtype=np.int32)
        nc, na = ctypes.c_int(), ctypes.c_long()
        self._chk(self.L.msc_cluster_labels_dd(self._ctx, rep, float(min_s),
                                               float("inf") if u_cut is None else float(u_cut),
                                               st.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), ctypes.byref(idx),
                                               labels.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                               ctypes.byref(nc), ctypes.byref(na)))
        return labels, nc.value, st, idx.value, na.value

    def work_counts(self, rep):
        c = (ctypes.c_long * 4)()
        self._chk(self.L.msc_work_counts(self._ctx, rep, c))
        return list(c)


def mt19937_state(seed):
    """MT19937 words after init_genrand(seed) and position 624, the state
    gsl_rng_set(ran, seed) leaves (the reference seeds ``ran`` with 1)."""
    mt = np.zeros(624, dtype=np.uint32)
    mt[0] = seed
    for i in range(1, 624):
        prev = int(mt[i - 1])
        mt[i] = (1812433253 * (prev ^ (prev >> 30)) + i) & 0xffffffff
    return mt, 624


def ini_state(N, L, length, dia, mass, temp, nmix=1000, mt=None, mt_index=624):
    """The reference's Signal=0 initial state (ini_pos, ``nmix`` GCAmix moves,
    ini_mom; v25/main.cpp:215-260) of N particles of one species, drawn from
    the generator state ``mt``/``mt_index`` (default: seeded 1, as the
    reference's ``ran``).  Returns a dict of r, p, l, ori (as set_state takes
    them) and the advanced generator (``mt``, ``mt_index``)."""
    lib = _lib.load()
    if mt is None:
        mt, mt_index = mt19937_state(1)
    mt = np.ascontiguousarray(mt, dtype=np.uint32).copy()
    idx = ctypes.c_int(mt_index)
    counts = np.array([N], dtype=np.int32)
    Lb = np.ascontiguousarray(L, dtype=np.float64)
    arr = [np.ascontiguousarray(np.broadcast_to(v, (N,)), dtype=np.float64) for v in (length, dia, mass)]
    r, p, l, ori = np.zeros((N, 3)), np.zeros((N, 3)), np.zeros((N, 3)), np.zeros((N, 3, 3))
    rc = lib.msc_ini_state(N, 1, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _d(Lb)[1],
                           *[_d(a)[1] for a in arr], temp, nmix, mt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                           ctypes.byref(idx), *[_d(a)[1] for a in (r, p, l, ori)])
    if rc == MSC_E_OVERLAP:
        raise Overlap(rc, "msc_ini_state: the lattice does not fit the box or its sites overlap")
    if rc:
        raise MscError(rc, "msc_ini_state failed")
    return {"r": r, "p": p, "l": l, "ori": ori, "mt": mt, "mt_index": idx.value}
