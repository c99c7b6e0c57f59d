"""ctypes binding of libmsc_hip.so (the C ABI declared in include/msc.h).

The library is built in-tree (``mdspherocyl_amd/libmsc_hip.so``) by
``__graft_entry__.build()``; importing this module never compiles anything and
fails loudly when the library is absent.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libmsc_hip.so")

MSC_OK, MSC_E_OVERLAP, MSC_E_DISTANCE, MSC_E_ARG, MSC_E_HIP, MSC_E_CAPACITY, MSC_E_STATE = range(7)
MSC_FP64, MSC_MIXED = 0, 1
MSC_ALL = -1
OBS_NAMES = ["K", "U", "P", "Uss", "Pss", "UdipR", "PdipR", "UdipK", "PdipK", "Mom2", "Ang2", "rebuild"]
OBS_N = len(OBS_NAMES)

# every symbol include/msc.h declares
EXPORTS = [
    "msc_create", "msc_destroy", "msc_last_error", "msc_version", "msc_set_state", "msc_get_state",
    "msc_get_forces", "msc_set_forces", "msc_get_thermostat", "msc_set_thermostat", "msc_neigh_list",
    "msc_get_neigh_list", "msc_check_list", "msc_force_sc", "msc_force_dipR", "msc_force_dipK",
    "msc_ext_field", "msc_self_ener", "msc_zero_forces", "msc_moveX", "msc_moveP", "msc_renorm",
    "msc_init_run", "msc_run", "msc_set_graph", "msc_cluster_labels", "msc_set_timing",
    "msc_kernel_time", "msc_reset_timing", "msc_work_counts", "msc_chunk_repeats", "msc_write_conf",
    "msc_read_conf", "msc_pair_correl", "msc_order_params", "msc_ini_state", "msc_set_contact_rows",
    "msc_pair_correl_noalloc", "msc_place_rods", "msc_cluster_labels_dd", "msc_get_tuning", "msc_set_tuning",
    "msc_autotune", "msc_set_neigh_list", "msc_order_params_ref",
]

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Tuning(ctypes.Structure):
    """msc_tuning: launch choices of a context."""
    _fields_ = [("graph", ctypes.c_int), ("fused_pairs", ctypes.c_int), ("pair_waves", ctypes.c_int),
                ("list_skin", ctypes.c_double), ("chunk", ctypes.c_int), ("overlap", ctypes.c_int),
                ("fuse_kick", ctypes.c_int)]

    def as_dict(self):
        return {"graph": self.graph, "fused_pairs": self.fused_pairs, "pair_waves": self.pair_waves,
                "list_skin": self.list_skin, "chunk": self.chunk, "overlap": self.overlap,
                "fuse_kick": self.fuse_kick}


class ReplicaDesc(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("nd", ctypes.c_int), ("L", ctypes.c_double * 3),
                ("alpha", ctypes.c_double), ("K", ctypes.c_int), ("sr_cut", ctypes.c_double),
                ("rc", ctypes.c_double), ("B", ctypes.c_double * 3), ("temp", ctypes.c_double),
                ("dt", ctypes.c_double), ("Q", ctypes.c_double), ("gamma", ctypes.c_double),
                ("s0", ctypes.c_double), ("dens", ctypes.c_double), ("dmin", ctypes.c_double),
                ("dmax", ctypes.c_double), ("thermo_seed", ctypes.c_ulong * 2)]


class MscError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("msc error %d: %s" % (code, msg))
        self.code = code


_lib = None


def load():
    """Load the in-tree HIP library (raises when it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError("libmsc_hip.so is not built; run __graft_entry__.build() first (%s)" % LIBPATH)
    L = ctypes.CDLL(LIBPATH)
    vp = ctypes.c_void_p
    sig = {
        "msc_create": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ReplicaDesc),
                        ctypes.POINTER(vp)], ctypes.c_int),
        "msc_destroy": ([vp], None),
        "msc_last_error": ([vp], ctypes.c_char_p),
        "msc_version": ([], ctypes.c_int),
        "msc_set_state": ([vp, ctypes.c_int] + [_dp] * 10, ctypes.c_int),
        "msc_get_state": ([vp, ctypes.c_int] + [_dp] * 4, ctypes.c_int),
        "msc_get_forces": ([vp, ctypes.c_int, _dp, _dp], ctypes.c_int),
        "msc_set_forces": ([vp, ctypes.c_int, _dp, _dp], ctypes.c_int),
        "msc_get_thermostat": ([vp, ctypes.c_int, _dp, _dp], ctypes.c_int),
        "msc_set_thermostat": ([vp, ctypes.c_int, _dp, _dp], ctypes.c_int),
        "msc_neigh_list": ([vp, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "msc_get_neigh_list": ([vp, ctypes.c_int, _ip, _ip, ctypes.c_long, ctypes.POINTER(ctypes.c_long)],
                               ctypes.c_int),
        "msc_set_neigh_list": ([vp, ctypes.c_int, ctypes.c_double, _ip, _ip], ctypes.c_int),
        "msc_check_list": ([vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _ip], ctypes.c_int),
        "msc_force_sc": ([vp, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp], ctypes.c_int),
        "msc_force_dipR": ([vp, ctypes.c_int, _dp, _dp], ctypes.c_int),
        "msc_force_dipK": ([vp, ctypes.c_int, _dp, _dp], ctypes.c_int),
        "msc_ext_field": ([vp, ctypes.c_int, _dp], ctypes.c_int),
        "msc_self_ener": ([vp, ctypes.c_int, _dp], ctypes.c_int),
        "msc_zero_forces": ([vp, ctypes.c_int], ctypes.c_int),
        "msc_moveX": ([vp, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "msc_moveP": ([vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp], ctypes.c_int),
        "msc_renorm": ([vp, ctypes.c_int], ctypes.c_int),
        "msc_init_run": ([vp], ctypes.c_int),
        "msc_run": ([vp, ctypes.c_int, ctypes.c_int, _dp], ctypes.c_int),
        "msc_set_graph": ([vp, ctypes.c_int], ctypes.c_int),
        "msc_get_tuning": ([vp, ctypes.POINTER(Tuning)], ctypes.c_int),
        "msc_set_tuning": ([vp, ctypes.POINTER(Tuning)], ctypes.c_int),
        "msc_autotune": ([vp, ctypes.c_int, ctypes.POINTER(Tuning), ctypes.POINTER(Tuning), _dp, ctypes.c_int, _ip],
                         ctypes.c_int),
        "msc_cluster_labels": ([vp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_uint),
                                _ip, _ip, _ip, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "msc_cluster_labels_dd": ([vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_uint),
                                   _ip, _ip, _ip, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "msc_set_contact_rows": ([vp, ctypes.c_int], ctypes.c_int),
        "msc_ini_state": ([ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int,
                           ctypes.POINTER(ctypes.c_uint), _ip, _dp, _dp, _dp, _dp], ctypes.c_int),
        "msc_order_params": ([vp, ctypes.c_int, _dp, _dp, _dp], ctypes.c_int),
        "msc_order_params_ref": ([vp, ctypes.c_int, _dp, _dp, _dp], ctypes.c_int),
        "msc_pair_correl": ([vp, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_double, ctypes.c_int, _dp,
                             ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "msc_pair_correl_noalloc": ([vp, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_double, ctypes.c_int, _dp,
                                     ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "msc_place_rods": ([ctypes.c_int, ctypes.c_double, _dp, ctypes.c_double, ctypes.c_ulonglong, _dp, _dp],
                           ctypes.c_int),
        "msc_set_timing": ([vp, ctypes.c_int], ctypes.c_int),
        "msc_kernel_time": ([vp, ctypes.c_char_p, _dp, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "msc_reset_timing": ([vp], ctypes.c_int),
        "msc_work_counts": ([vp, ctypes.c_int, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "msc_chunk_repeats": ([vp, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "msc_write_conf": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int] + [_dp] * 9 + [_ip, _dp], ctypes.c_int),
        "msc_read_conf": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int] + [_dp] * 9 + [_ip, _dp], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L
