"""The reference's run files: ``input.txt`` (r_input, v25/functions.cpp:10-87)
and ``conf.dat`` (print_conf / chrg_conf, v25/functions.cpp:1181-1296).

``read_input`` follows r_input token for token (labels are skipped by
position, as the reference's stream does), ``replica_desc_from_input`` turns
its result into a replica of the library (box from the number density as
v25/main.cpp:216-219 does), and ``read_conf`` / ``write_conf`` call the
library's host-only conf.dat reader and writer (``msc_read_conf`` /
``msc_write_conf``), whose output is byte-identical to print_conf's.
"""
import ctypes
import math

import numpy as np

from . import _lib


def _tokens(path):
    with open(path) as f:
        return f.read().split()


def read_input(path):
    """The values r_input reads from ``path``, keyed by their reference
    names; ``dip`` holds per species a list of (position, orientation[3],
    moment) of its dipoles."""
    t = _tokens(path)
    pos = [0]

    def skip(n):
        pos[0] += n

    def num(kind=float):
        v = kind(t[pos[0]])
        pos[0] += 1
        return v

    r = {}
    skip(3); r["Nsteps"] = num(int)
    skip(3); r["pre_steps"] = num(int)
    skip(1); r["signal"] = num(int)
    skip(3); n_comp = num(int)
    r["n_comp"] = n_comp
    skip(3); r["n_part"] = [num(int) for _ in range(n_comp)]
    skip(1); r["length"] = [num() for _ in range(n_comp)]
    skip(1); r["diameter"] = [num() for _ in range(n_comp)]
    skip(1); r["mass"] = [num() for _ in range(n_comp)]
    skip(2); r["B"] = [num() for _ in range(3)]
    skip(1); r["dens"] = num()
    skip(1); r["temp"] = num()
    skip(1); r["Q"] = num()
    skip(1); r["gamma"] = num()
    skip(5); r["ArLx"] = num(); r["ArLy"] = num()
    skip(3); r["Dt"] = num()
    skip(2); r["sr_cut"] = num()
    skip(3); r["K_max"] = num(int)
    skip(2); r["alpha"] = num()
    skip(3); r["bar_w"] = num()
    skip(3); r["corrf"] = num(int)
    skip(2); r["tcfd"] = num(int)
    skip(2); r["tcfv"] = num(int)
    skip(2); r["cl_fr"] = num(int)
    skip(2); r["min_s"] = num()
    skip(2); r["iprint"] = num(int)
    skip(1)
    skip(5)
    dip = []
    for _ in range(n_comp):
        skip(1)
        ndp = num(int)
        ds = []
        for _ in range(ndp):
            p = num()
            skip(1)
            o = [num() for _ in range(3)]
            skip(1)
            mu = num()
            ds.append((p, o, mu))
        dip.append(ds)
    r["dip"] = dip
    return r


def box_from_input(inp):
    """Box edges from the number density (v25/main.cpp:216-219)."""
    n = sum(inp["n_part"])
    Lz = math.pow(float(n) / (inp["ArLx"] * inp["ArLy"] * inp["dens"]), 1.0 / 3)
    return (inp["ArLx"] * Lz, inp["ArLy"] * Lz, Lz)


def replica_desc_from_input(inp, box=None, **over):
    """A replica of the reference run described by ``inp`` (reference mode:
    every minimum-image real-space pair, ``rc = 0``).  ``box`` defaults to
    the one the number density gives (a new run, signal 0); a continued run
    (signal 1) takes the box line of its conf.dat, as chrg_conf does
    (v25/main.cpp:262-267).  Keyword arguments override fields of
    ``replica_desc``."""
    from .system import replica_desc
    nd = max((len(d) for d in inp["dip"]), default=1)
    kw = dict(rc=0.0, nd=max(nd, 1), B=tuple(inp["B"]), temp=inp["temp"], dt=inp["Dt"], Q=inp["Q"],
              gamma=inp["gamma"], dens=inp["dens"])
    kw.update(over)
    L = box_from_input(inp) if box is None else tuple(box)
    return replica_desc(sum(inp["n_part"]), L, inp["alpha"], inp["K_max"], inp["sr_cut"], **kw)


def _ptr(a, kind=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(kind)) if a is not None else None


def read_conf(path, N, nd):
    """conf.dat of N particles with nd dipoles each: dict of L, r, p, ori
    (N x 3 x 3), l, mass, dpos, dori, dmu, clust and thermo (s0, strn0, s1,
    strn1)."""
    L = _lib.load()
    c = {"L": np.zeros(3), "r": np.zeros((N, 3)), "p": np.zeros((N, 3)), "ori": np.zeros((N, 3, 3)),
         "l": np.zeros((N, 3)), "mass": np.zeros(N), "dpos": np.zeros((N, nd, 3)), "dori": np.zeros((N, nd, 3)),
         "dmu": np.zeros((N, nd)), "clust": np.zeros(N, dtype=np.int32), "thermo": np.zeros(4)}
    rc = L.msc_read_conf(str(path).encode(), N, nd, _ptr(c["L"]), _ptr(c["r"]), _ptr(c["p"]), _ptr(c["ori"]),
                         _ptr(c["l"]), _ptr(c["mass"]), _ptr(c["dpos"]), _ptr(c["dori"]), _ptr(c["dmu"]),
                         _ptr(c["clust"], ctypes.c_int), _ptr(c["thermo"]))
    if rc:
        raise _lib.MscError(rc, "msc_read_conf(%s) failed" % path)
    return c


def write_conf(path, c):
    """Write a dict as read_conf returns (clust may be missing: 0)."""
    L = _lib.load()
    N, nd = c["r"].shape[0], np.asarray(c["dmu"]).reshape(c["r"].shape[0], -1).shape[1]
    arr = {k: np.ascontiguousarray(c[k], dtype=np.float64)
           for k in ("L", "r", "p", "ori", "l", "mass", "dpos", "dori", "dmu", "thermo")}
    cl = np.ascontiguousarray(c["clust"], dtype=np.int32) if c.get("clust") is not None else None
    rc = L.msc_write_conf(str(path).encode(), N, nd, _ptr(arr["L"]), _ptr(arr["r"]), _ptr(arr["p"]),
                          _ptr(arr["ori"]), _ptr(arr["l"]), _ptr(arr["mass"]), _ptr(arr["dpos"]),
                          _ptr(arr["dori"]), _ptr(arr["dmu"]), _ptr(cl, ctypes.c_int), _ptr(arr["thermo"]))
    if rc:
        raise _lib.MscError(rc, "msc_write_conf(%s) failed" % path)


def species_arrays(inp):
    """Per-particle length, diameter and dipole body-frame description
    (dpc: signed distance along the axis, doc: orientation in body axes, dmu:
    moment) of the run in ``inp``, species-major as the reference orders its
    particles (v25/main.cpp:165-177 builds them from the same fields).  Every
    species must carry the same number of dipoles."""
    nds = {len(d) for d in inp["dip"]}
    if len(nds) != 1:
        raise ValueError("species with different dipole counts are not supported")
    nd = nds.pop()
    N = sum(inp["n_part"])
    length, dia = np.zeros(N), np.zeros(N)
    dpc, doc, dmu = np.zeros((N, max(nd, 1))), np.zeros((N, max(nd, 1), 3)), np.zeros((N, max(nd, 1)))
    if nd == 0:
        doc[:, 0, 2] = 1.0
    g = 0
    for s, n in enumerate(inp["n_part"]):
        length[g:g + n], dia[g:g + n] = inp["length"][s], inp["diameter"][s]
        for w, (pos, o, mu) in enumerate(inp["dip"][s]):
            dpc[g:g + n, w], doc[g:g + n, w], dmu[g:g + n, w] = pos, o, mu
        g += n
    return length, dia, dpc, doc, dmu


def lab_dipoles(r, ori, dpc, doc):
    """Lab-frame dipole positions and orientations as Spherocyl::set_dip
    forms them (v25/classes.cpp:160-170): pos = r + dpc * e, e the third row
    of the orientation matrix, and ori = transpose(ori) * doc."""
    e = ori[:, 2, :]
    dpos = r[:, None, :] + dpc[:, :, None] * e[:, None, :]
    dori = np.einsum("gji,gwj->gwi", ori, doc)
    return dpos, dori


def load_conf(sysm, rep, c, inp):
    """Put a conf.dat state (as read_conf returns it) into replica ``rep``
    of ``sysm`` the way chrg_conf does (v25/functions.cpp:1240-1296):
    positions, momenta, orientations, angular momenta, masses and both
    thermostats come from the file; the dipoles' lab-frame columns are
    skipped and rebuilt from the body-frame description in ``inp``."""
    length, dia, dpc, doc, dmu = species_arrays(inp)
    sysm.set_state(rep, c["r"], c["p"], c["l"], c["ori"], length, dia, c["mass"], dpc, doc, dmu)
    th = np.asarray(c["thermo"], dtype=np.float64)
    sysm.set_thermostat(rep, th[0::2], th[1::2])


def conf_from_system(sysm, rep, inp, mass, clust=None):
    """The conf.dat of replica ``rep`` as print_conf writes it
    (v25/functions.cpp:1181-1225): box, state read back from the device,
    lab-frame dipoles, cluster labels (0 when none are given) and the two
    thermostats."""
    r, p, l, ori = sysm.get_state(rep)
    length, dia, dpc, doc, dmu = species_arrays(inp)
    dpos, dori = lab_dipoles(r, ori, dpc, doc)
    s, strn = sysm.thermostat(rep)
    N = r.shape[0]
    return {"L": np.asarray(sysm.descs[rep].L[:3], dtype=np.float64), "r": r, "p": p, "ori": ori, "l": l,
            "mass": np.asarray(mass, dtype=np.float64).reshape(N), "dpos": dpos, "dori": dori, "dmu": dmu,
            "clust": np.zeros(N, dtype=np.int32) if clust is None else np.asarray(clust, dtype=np.int32),
            "thermo": np.array([s[0], strn[0], s[1], strn[1]])}


def correl_bars(L, bar_w):
    """Number of histogram bars the reference allocates (v25/main.cpp:251):
    floor(sqrt(3) * Lmin / (2 * bar_w))."""
    return int(math.floor(math.sqrt(3.) * min(L) / (2. * bar_w)))


def correl_sets(nspec):
    """Histogram sets of ``nspec`` species: one per species (pairs within
    it), then one per species pair m < n in row-major order."""
    return nspec + nspec * (nspec - 1) // 2


def _fmt_line_cor(first, v):
    # the first value goes out in the stream's default format; setprecision(8)
    # and fixed stay set for every later line (v25/functions.cpp:2491-2495)
    return ("%g" % v) if first else ("%.8f" % v)


def correlation_text(hist, n_part, Lmin, bar_w, cor_cont):
    """correlation.dat as nor_correl writes it (v25/functions.cpp:2468-2538)
    from accumulated histograms ``hist`` [sets][4][n_bars] (sets as
    correl_sets orders them), the species counts, the smallest box edge, the
    bar width and cor_cont (the sum of 1/V over accumulated frames)."""
    NC = len(n_part)
    hist = np.asarray(hist, dtype=np.float64).reshape(correl_sets(NC), 4, -1)
    n_bars = hist.shape[2]
    pd = 4. * math.pi
    bar3 = bar_w * bar_w * bar_w / 12.
    sq2 = math.sqrt(2.)
    out = []
    for b in range(n_bars):
        pos = (float(b) + 0.5) * bar_w
        if pos >= Lmin / 2. and pos <= Lmin / sq2:
            norm = pd * pos * pos * (3 * Lmin / (2. * pos) - 2.) * bar_w
        else:
            norm = pd * (pos * pos * bar_w + bar3)
        nnorm = norm * cor_cont
        s = [_fmt_line_cor(b == 0, cor_cont), " %.8f %.4f & " % (norm, pos)]
        for n in range(NC):
            NN = n_part[n]
            NN1 = NN - 1
            if NN1 > 0:
                h = hist[n, :, b]
                s.append("%.8f " % (1. * h[0] / (nnorm * NN * NN1)))
                s.append("%.8f " % (3. * h[1] / (nnorm * NN * NN1)))
                s.append("%.8f " % ((3. / 2) * h[2] / (nnorm * NN * NN1)))
                s.append("%.8f " % (5. * h[3] / (nnorm * NN * NN1)))
            else:
                s.append("0 ")
            s.append("| ")
        cc = NC
        for m in range(NC - 1):
            MM = n_part[m]
            for n in range(m + 1, NC):
                NN = n_part[n]
                h = hist[cc, :, b]
                s.append("%.8f " % (1. * h[0] / (nnorm * NN * MM)))
                s.append("%.8f " % (3. * h[1] / (nnorm * NN * MM)))
                s.append("%.8f " % ((3. / 2) * h[2] / (nnorm * NN * MM)))
                s.append("%.8f " % (5. * h[3] / (nnorm * NN * MM)))
                s.append("| ")
                cc += 1
        out.append("".join(s) + "\n")
    return "".join(out)
