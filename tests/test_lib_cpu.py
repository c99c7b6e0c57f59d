"""CPU-side checks of the HIP library: it loads, exports every symbol of
include/msc.h, and rejects bad arguments before touching a GPU."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "msc.h")).read()
    return sorted(set(re.findall(r"\b(msc_[a-z_A-Z]+)\s*\(", text)))


def test_header_and_binding_agree():
    import mdspherocyl_amd as m
    assert header_symbols() == sorted(m.EXPORTS)


def test_library_exports_every_declared_symbol():
    import mdspherocyl_amd as m
    if not os.path.exists(m.LIBPATH):
        pytest.skip("library not built")
    out = subprocess.run(["nm", "-D", "--defined-only", m.LIBPATH], capture_output=True, text=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in header_symbols() if s not in names]
    assert not missing, missing
    L = m.load()
    assert L.msc_version() == 1


def test_bad_arguments_rejected_without_a_device():
    import mdspherocyl_amd as m
    if not os.path.exists(m.LIBPATH):
        pytest.skip("library not built")
    L = m.load()
    ctx = ctypes.c_void_p()
    assert L.msc_create(0, 0, 0, None, ctypes.byref(ctx)) == m._lib.MSC_E_ARG
    d = m.replica_desc(10, 5.0, 0.5, 1, 4.0)  # K < 2 is invalid
    arr = (m.ReplicaDesc * 1)(d)
    assert L.msc_create(0, 0, 1, arr, ctypes.byref(ctx)) == m._lib.MSC_E_ARG
    assert L.msc_last_error(None) == b"no context"


def test_new_entry_points_reject_bad_arguments_without_a_device(tmp_path):
    """The correlation, order-parameter, conf.dat and initial-state entry
    points refuse missing contexts, pointers and sizes before any device
    work; the conf.dat reader reports a missing file."""
    import mdspherocyl_amd as m
    if not os.path.exists(m.LIBPATH):
        pytest.skip("library not built")
    L = m.load()
    E = m._lib.MSC_E_ARG
    h = (ctypes.c_double * 8)()
    cnt = (ctypes.c_int * 1)(10)
    assert L.msc_pair_correl(None, 0, 1, cnt, 0.1, 2, h, None) == E
    g1, g2, M = ctypes.c_double(), ctypes.c_double(), (ctypes.c_double * 3)()
    assert L.msc_order_params(None, 0, ctypes.byref(g1), ctypes.byref(g2), M) == E
    assert L.msc_order_params_ref(None, 0, ctypes.byref(g1), ctypes.byref(g2), M) == E
    z = (ctypes.c_double * 64)()
    assert L.msc_read_conf(str(tmp_path / "missing.dat").encode(), 1, 1, z, z, z, z, z, z, z, z, z, None, z) == E
    assert L.msc_write_conf(None, 1, 1, z, z, z, z, z, z, z, z, z, None, z) == E
    mt = (ctypes.c_uint * 624)()
    idx = ctypes.c_int(624)
    one = (ctypes.c_double * 3)(1.0, 1.0, 1.0)
    # N does not match the species counts
    assert L.msc_ini_state(2, 1, cnt, one, one, one, one, 1.0, 0, mt, ctypes.byref(idx), z, z, z, z) == E
    # generator position outside [0, 624]
    idx.value = 700
    assert L.msc_ini_state(10, 1, cnt, one, z, z, z, 1.0, 0, mt, ctypes.byref(idx), z, z, z, z) == E
