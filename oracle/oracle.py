"""ctypes front end of the oracle (test infrastructure only).

The oracle is the CPU FP64 restatement of the reference's force path in
``oracle/msc_oracle.c``.  Only tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module; the product package
``mdspherocyl_amd`` never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libmsc_oracle.so")
LIBPATH_OMP = os.path.join(HERE, "libmsc_oracle_omp.so")
MAXSPEC = 8
OBS_NAMES = ["K", "U", "P", "Uss", "Pss", "UdipR", "PdipR", "UdipK", "PdipK",
             "Mom2", "Ang2", "rebuild"]
OBS_N = len(OBS_NAMES)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class COsys(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("nd", ctypes.c_int), ("nspec", ctypes.c_int),
                ("counts", ctypes.c_int * MAXSPEC), ("L", ctypes.c_double * 3),
                ("r", _dp), ("p", _dp), ("l", _dp), ("ori", _dp), ("F", _dp), ("T", _dp),
                ("len", _dp), ("dia", _dp), ("mass", _dp), ("inert", _dp),
                ("dpc", _dp), ("doc", _dp), ("dmu", _dp), ("dr", _dp), ("du", _dp),
                ("clust", _ip)]


class COmt(ctypes.Structure):
    _fields_ = [("mt", ctypes.c_uint32 * 624), ("mti", ctypes.c_int)]


class COtherm(ctypes.Structure):
    _fields_ = [("Q", ctypes.c_double), ("s", ctypes.c_double), ("g", ctypes.c_double),
                ("strn", ctypes.c_double), ("ran", COmt)]


class CRunParams(ctypes.Structure):
    _fields_ = [("dens", ctypes.c_double), ("temp", ctypes.c_double), ("dt", ctypes.c_double),
                ("sr_cut", ctypes.c_double), ("alpha", ctypes.c_double), ("min_s", ctypes.c_double),
                ("Q", ctypes.c_double), ("gamma", ctypes.c_double), ("s0", ctypes.c_double),
                ("B", ctypes.c_double * 3), ("K", ctypes.c_int), ("cl_fr", ctypes.c_int),
                ("rc_real", ctypes.c_double), ("dipK_on", ctypes.c_int), ("dipR_on", ctypes.c_int)]


class CRun(ctypes.Structure):
    _fields_ = [("Tab_alpha", ctypes.c_double), ("Tab", _dp), ("Kvec", _dp),
                ("off", _ip), ("nbr", _ip), ("cap", ctypes.c_long), ("snap", _dp),
                ("r_list_min", ctypes.c_double), ("r_list_max", ctypes.c_double),
                ("Uself", ctypes.c_double), ("dmax", ctypes.c_double), ("dmin", ctypes.c_double),
                ("th", COtherm * 2), ("ran", COmt), ("PT", ctypes.c_double * 9)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    """The serial oracle: the checker."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        _lib = _bind(ctypes.CDLL(LIBPATH))
    return _lib


_lib_omp = None


def lib_omp():
    """The same routines built with OpenMP (force loops split over threads):
    only for timing the host baseline; its sums are taken in another order."""
    global _lib_omp
    if _lib_omp is None:
        if not os.path.exists(LIBPATH_OMP):
            build()
        _lib_omp = _bind(ctypes.CDLL(LIBPATH_OMP))
    return _lib_omp


def _bind(L):
    """Argument and result types of the exported routines."""
    P = ctypes.POINTER
    L.o_segdist.restype = ctypes.c_double
    L.o_segdist.argtypes = [_dp, _dp, ctypes.c_double, _dp, _dp, ctypes.c_double,
                            _dp, _dp, _dp, _dp]
    L.o_spherocyl_dist.restype = ctypes.c_double
    L.o_neigh_list.restype = ctypes.c_long
    L.o_neigh_list.argtypes = [P(COsys), ctypes.c_double, _ip, _ip, ctypes.c_long]
    L.o_round12.restype = ctypes.c_double
    L.o_round12.argtypes = [ctypes.c_double]
    L.o_check_list.argtypes = [P(COsys), _dp, ctypes.c_double, ctypes.c_double]
    L.o_force_sc.argtypes = [P(COsys), ctypes.c_double, _ip, _ip, _dp, _dp, _dp]
    L.o_tabul.argtypes = [_dp, ctypes.c_double]
    L.o_force_dipR.argtypes = [P(COsys), _dp, ctypes.c_double, _dp, _dp]
    L.o_wave_v.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_int]
    L.o_force_dipK.argtypes = [P(COsys), ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
    L.o_ext_field.argtypes = [P(COsys), _dp, _dp]
    L.o_self_ener.restype = ctypes.c_double
    L.o_self_ener.argtypes = [P(COsys)]
    L.o_set_inertia.argtypes = [P(COsys), ctypes.c_int]
    L.o_set_dip.argtypes = [P(COsys), ctypes.c_int]
    L.o_moveX.argtypes = [P(COsys), ctypes.c_double]
    L.o_renorm.argtypes = [P(COsys), ctypes.c_int]
    L.omt_seed.argtypes = [P(COmt), ctypes.c_ulong]
    L.omt_u32.restype = ctypes.c_uint32
    L.omt_u32.argtypes = [P(COmt)]
    L.omt_uniform.restype = ctypes.c_double
    L.omt_uniform.argtypes = [P(COmt)]
    L.o_gauss.argtypes = [ctypes.c_double, _dp, _dp]
    L.o_ini_pos.argtypes = [P(COsys)]
    L.o_gcamix.argtypes = [P(COsys), P(COmt)]
    L.o_ini_mom.argtypes = [P(COsys), ctypes.c_double, P(COmt)]
    L.o_dist_test.argtypes = [P(COsys), ctypes.c_int, ctypes.c_int, ctypes.c_double]
    L.o_clust_conf.argtypes = [P(COsys), ctypes.c_double, P(COmt)]
    L.o_clust_conf_dd.argtypes = [P(COsys), ctypes.c_double, ctypes.c_double, P(COmt)]
    L.o_udd.restype = ctypes.c_double
    L.o_udd.argtypes = [P(COsys), ctypes.c_int, ctypes.c_int]
    L.o_contact_pairs.restype = ctypes.c_long
    L.o_contact_pairs.argtypes = [P(COsys), ctypes.c_long, _ip, _ip, ctypes.c_double, ctypes.c_double,
                                  ctypes.POINTER(ctypes.c_ubyte)]
    L.o_acc_correl.argtypes = [P(COsys), ctypes.c_double, ctypes.c_int, _dp]
    L.o_acc_correl.restype = ctypes.c_long
    _fp = ctypes.POINTER(ctypes.c_float)
    L.o_jacobi3.argtypes = [_fp, _fp, _fp]
    L.o_oo_params.argtypes = [P(COsys), _dp, _dp, _dp]
    L.o_run_init.argtypes = [P(COsys), P(CRunParams), P(CRun), ctypes.c_int]
    L.o_run_steps.argtypes = [P(COsys), P(CRunParams), P(CRun), ctypes.c_int, ctypes.c_int, _dp]
    L.o_run_free.argtypes = [P(CRun)]
    L.o_forces.argtypes = [P(COsys), P(CRunParams), P(CRun), _dp]
    L.o_threads.restype = ctypes.c_int
    L.o_threads.argtypes = []
    return L


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


class OSys:
    """Particle system in the oracle's SoA layout (species-major order).

    ``counts`` lists the particles of each species; per-particle shape arrays
    may differ within a species (the polydisperse extension), but the
    reference's per-species routines (self_ener, GCAmix radii, ini_mom) read
    the first particle of a species, as the reference does."""

    def __init__(self, N, nd=1, L=(1.0, 1.0, 1.0), counts=None):
        self.N, self.nd = int(N), int(nd)
        self.counts = list(counts) if counts is not None else [self.N]
        assert sum(self.counts) == self.N and len(self.counts) <= MAXSPEC
        self.L = np.array(L, dtype=np.float64)
        z = lambda *s: np.zeros(s, dtype=np.float64)
        self.r, self.p, self.l = z(N, 3), z(N, 3), z(N, 3)
        self.ori = np.tile(np.eye(3), (N, 1, 1)).astype(np.float64)
        self.F, self.T = z(N, 3), z(N, 3)
        self.len, self.dia, self.mass = z(N), np.ones(N), np.ones(N)
        self.inert = z(N, 3)
        self.dpc = z(N, nd)
        self.doc = z(N, nd, 3)
        self.doc[:, :, 2] = 1.0
        self.dmu = z(N, nd)
        self.dr, self.du = z(N, nd, 3), z(N, nd, 3)
        self.clust = np.zeros(N, dtype=np.int32)
        self._c = None

    def c(self):
        s = COsys()
        s.N, s.nd, s.nspec = self.N, self.nd, len(self.counts)
        for i, v in enumerate(self.counts):
            s.counts[i] = v
        for a in range(3):
            s.L[a] = self.L[a]
        for name in ["r", "p", "l", "ori", "F", "T", "len", "dia", "mass", "inert",
                     "dpc", "doc", "dmu", "dr", "du"]:
            setattr(s, name, dptr(getattr(self, name)))
        s.clust = iptr(self.clust)
        self._c = s
        return ctypes.byref(s)

    def set_inertia(self):
        c = self.c()
        for i in range(self.N):
            lib().o_set_inertia(c, i)

    def set_dip(self):
        c = self.c()
        for i in range(self.N):
            lib().o_set_dip(c, i)

    def axes(self):
        return self.ori[:, 2, :].copy()

    def copy(self):
        o = OSys(self.N, self.nd, self.L, self.counts)
        for k, v in self.__dict__.items():
            if isinstance(v, np.ndarray):
                setattr(o, k, v.copy())
        return o


def run_params(**kw):
    P = CRunParams()
    defaults = dict(dens=0.1, temp=1.0, dt=0.002, sr_cut=5.0, alpha=0.3, min_s=0.2, Q=5.0,
                    gamma=0.0, s0=0.0, B=(0.0, 0.0, 0.0), K=5, cl_fr=0, rc_real=0.0,
                    dipK_on=1, dipR_on=1)
    defaults.update(kw)
    for k, v in defaults.items():
        if k == "B":
            for a in range(3):
                P.B[a] = v[a]
        else:
            setattr(P, k, v)
    return P


def osys_from_replica(desc, st):
    """Oracle copy of a replica (descriptor and state of
    mdspherocyl_amd.configs)."""
    S = OSys(desc.N, 1, L=tuple(desc.L))
    for k, key in (("r", "r"), ("p", "p"), ("l", "l"), ("ori", "ori"), ("len", "length"), ("dia", "dia"),
                   ("mass", "mass"), ("dpc", "dpc"), ("doc", "doc"), ("dmu", "dmu")):
        getattr(S, k)[:] = st[key]
    S.set_inertia()
    S.set_dip()
    return S


def run_params_from_replica(desc):
    """Run parameters of a replica descriptor (r_c > 0: the cut-off mode)."""
    return run_params(dens=desc.dens, temp=desc.temp, dt=desc.dt, sr_cut=desc.sr_cut, alpha=desc.alpha,
                      Q=desc.Q, gamma=desc.gamma, s0=desc.s0, B=tuple(desc.B), K=desc.K, rc_real=desc.rc)


class ORun:
    """A reference-style run (v25/main.cpp) driven through the oracle."""

    def __init__(self, sysobj, params, seed_init):
        self.S, self.P = sysobj, params
        self.R = CRun()
        rc = lib().o_run_init(self.S.c(), ctypes.byref(self.P), ctypes.byref(self.R),
                              int(bool(seed_init)))
        if rc:
            raise RuntimeError("oracle init failed with code %d" % rc)
        self.step = 0

    def run(self, nsteps):
        obs = np.zeros((nsteps, OBS_N), dtype=np.float64)
        rc = lib().o_run_steps(self.S.c(), ctypes.byref(self.P), ctypes.byref(self.R),
                               self.step, nsteps, dptr(obs))
        if rc:
            raise RuntimeError("oracle step failed with code %d" % rc)
        self.step += nsteps
        return obs

    def neighbour_list(self):
        N = self.S.N
        off = np.ctypeslib.as_array(self.R.off, shape=(N + 1,)).copy()
        nbr = np.ctypeslib.as_array(self.R.nbr, shape=(int(off[-1]),)).copy() if off[-1] else np.zeros(0, np.int32)
        return off, nbr

    def free(self):
        lib().o_run_free(ctypes.byref(self.R))
