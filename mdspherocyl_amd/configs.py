"""Synthetic systems of the benchmark workloads.

Each workload is a list of replicas; a replica is a parameter set for
``replica_desc`` plus an initial state with no two rods closer than one
diameter: c1 puts centres on a simple-cubic lattice with axes uniform on the
sphere; c2-c5 use random sequential addition of uniform centres and axes (c5, with
a quarter of a million rods, through the library's host routine
msc_place_rods)
(random axes on any lattice jam at phi >= 0.15 for these lengths: even rods
aligned along the cube diagonal of a simple-cubic lattice at phi = 0.2 touch
once L > 3.2); replicas denser than phi = 0.22, where addition slows down
near jamming, start as an orthorhombic lattice of rods along +-z with random
polarity and a 0.02 rad random tilt (the reference's own ini_pos start is an
aligned lattice too, v25/functions.cpp:491-584).  Maxwell momenta at T, zero
total momentum.  Randomly added rods are numbered along a Morton curve
(neighbours in space are neighbours in index).  One axial point dipole
per particle at the centre.  Lengths are Gaussian, truncated at +-3 sd, and
the box follows from the packing fraction phi = sum(v_i) / V with
v = pi D^2 L / 4 + pi D^3 / 6.

Workloads (reference "No. of Kvectors" K = k_max + 1):
  c1  N=256, L=3, phi=0.10, mu^2=2, alpha=5.6/L_box, every minimum-image
      real-space pair, K=6
  c2  N=4096, L~N(3,0.3), phi=0.20, mu^2=4, alpha=7/L_box, every
      minimum-image pair, K=8
  c3  N=32768, L~N(3,0.3), phi=0.20, mu^2=4, r_c=3.5, alpha=3/r_c, K=11
  c4  64 replicas, N_i~U{512..8192}, phi_i~U[0.05,0.30], mu_i^2~U[1,6],
      r_c=3.5, alpha=3/r_c, K=8
  c5  N=262144, L~N(5,0.5), phi=0.15, mu^2=6, r_c=3.5, alpha=3/r_c, K=13
All: D=1, m=1, T=1, dt=0.002, Q=5, gamma=0 (Nose-Hoover), seed = workload number.
"""
import math

import numpy as np

from .system import replica_desc

WORKLOADS = ("c1", "c2", "c3", "c4", "c5")


def segdist2(rA, uA, LA, rB, uB, LB, box):
    """Squared minimum distance of segment pairs (vectorised, clamped
    closest-point method); used only to place the initial state."""
    d = rA - rB
    d -= box * np.floor(d / box + 0.5)
    d1 = uA * LA[:, None]
    d2 = uB * LB[:, None]
    r = d - 0.5 * d1 + 0.5 * d2
    a = np.einsum("ij,ij->i", d1, d1)
    e = np.einsum("ij,ij->i", d2, d2)
    f = np.einsum("ij,ij->i", d2, r)
    b = np.einsum("ij,ij->i", d1, d2)
    c = np.einsum("ij,ij->i", d1, r)
    den = a * e - b * b
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.where(den > 1e-12, np.clip((b * f - c * e) / den, 0, 1), 0.0)
        t = np.where(e > 0, (b * s + f) / e, 0.0)
        lo, hi = t < 0, t > 1
        s = np.where(lo, np.where(a > 0, np.clip(-c / a, 0, 1), 0.0), s)
        s = np.where(hi, np.where(a > 0, np.clip((b - c) / a, 0, 1), 0.0), s)
        t = np.clip(t, 0, 1)
    v = r + s[:, None] * d1 - t[:, None] * d2
    return np.einsum("ij,ij->i", v, v)


def _centres_within(ra, rb, reach, box):
    """Minimum-image centre distance below reach (vectorised)."""
    d = ra - rb
    d -= box * np.floor(d / box + 0.5)
    return np.einsum("ij,ij->i", d, d) < reach * reach


def frames(u):
    """Orientation matrices with row 2 = u (rows = body axes)."""
    u = u / np.linalg.norm(u, axis=1)[:, None]
    a = np.where(np.abs(u[:, :1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
    e1 = np.cross(a, u)
    e1 /= np.linalg.norm(e1, axis=1)[:, None]
    e0 = np.cross(e1, u)
    return np.stack([e0, e1, u], axis=1)


def random_units(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def fisher_axes(rng, n, kappa):
    """Axes about +-z with von Mises-Fisher concentration kappa (random sign)."""
    w = 1.0 + np.log(rng.uniform(np.exp(-2 * kappa), 1.0, size=n)) / kappa
    phi = rng.uniform(0, 2 * np.pi, size=n)
    s = np.sqrt(np.clip(1 - w * w, 0, 1))
    v = np.stack([s * np.cos(phi), s * np.sin(phi), w], axis=1)
    return v * rng.choice([-1.0, 1.0], size=(n, 1))


def truncated_normal(rng, n, mean, sd, width=3.0):
    out = rng.normal(mean, sd, size=n)
    bad = np.abs(out - mean) > width * sd
    while bad.any():
        out[bad] = rng.normal(mean, sd, size=bad.sum())
        bad = np.abs(out - mean) > width * sd
    return out


def place_lattice(rng, N, lengths, box, dmin=1.0, batch=128, max_batches=64):
    """Simple-cubic lattice centres with random axes, pairwise segment
    distance >= dmin: rods are visited in lattice order and each keeps the
    first random axis clear of the rods already placed around it."""
    n = int(math.ceil(N ** (1.0 / 3) - 1e-9))
    a = box / n
    idx = np.arange(n ** 3)[:N]
    coords = np.stack([idx % n, (idx // n) % n, idx // (n * n)], axis=1)
    r = coords * a + 0.5 * a - 0.5 * box
    u = np.zeros((N, 3))
    reach = lengths.max() + dmin
    m = int(math.ceil(reach / a))
    offs = np.array([(i, j, k) for i in range(-m, m + 1) for j in range(-m, m + 1) for k in range(-m, m + 1)
                     if (i, j, k) != (0, 0, 0) and (max(0, abs(i) - 1) ** 2 + max(0, abs(j) - 1) ** 2 +
                                                    max(0, abs(k) - 1) ** 2) * a * a < reach * reach])
    site = -np.ones(n ** 3, dtype=np.int64)
    site[idx] = np.arange(N)
    for i in range(N):
        c2 = (coords[i] + offs) % n
        nb = site[c2[:, 0] + n * c2[:, 1] + n * n * c2[:, 2]]
        nb = nb[(nb >= 0) & (nb < i)]
        for _ in range(max_batches):
            cand = random_units(rng, batch)
            if nb.size == 0:
                u[i] = cand[0]
                break
            k = nb.size
            d2 = segdist2(np.repeat(r[i:i + 1], batch * k, 0), np.repeat(cand, k, 0),
                          np.full(batch * k, lengths[i]), np.tile(r[nb], (batch, 1)), np.tile(u[nb], (batch, 1)),
                          np.tile(lengths[nb], batch), box)
            ok = (d2.reshape(batch, k) >= dmin * dmin).all(axis=1)
            if ok.any():
                u[i] = cand[np.argmax(ok)]
                break
        else:
            raise RuntimeError("could not place rod %d of %d without contact" % (i, N))
    return r, u


def place_rsa(rng, N, lengths, box, dmin=1.0, batch=512, max_batches=200000):
    """Random sequential addition: uniform centres and axes, each candidate
    kept when its segment distance to every rod already in the box is >= dmin.
    Should acceptance fall below 2 %, axes are drawn ever closer to +-z
    (von Mises-Fisher) so that dense replicas still fill."""
    reach = lengths.max() + dmin
    nc = max(1, int(box // reach))
    cw = box / nc
    cap = 8
    members = -np.ones((nc ** 3, cap), dtype=np.int64)
    fill = np.zeros(nc ** 3, dtype=np.int64)
    r = np.zeros((N, 3))
    u = np.zeros((N, 3))
    offs = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)])
    if nc < 3:
        offs = np.zeros((1, 3), dtype=np.int64)
    placed = 0
    kappa, tried, got = 0.0, 0, 0
    for _ in range(max_batches):
        if placed == N:
            break
        B = batch
        cr = rng.uniform(-0.5 * box, 0.5 * box, size=(B, 3))
        cu = random_units(rng, B) if kappa == 0 else fisher_axes(rng, B, kappa)
        cl = lengths[placed:placed + B]
        if cl.size < B:
            cl = np.concatenate([cl, np.full(B - cl.size, lengths[placed])])
        cc = np.minimum(((cr + 0.5 * box) // cw).astype(np.int64), nc - 1)
        nbc = (cc[:, None, :] + offs[None, :, :]) % nc
        cid = nbc[:, :, 0] + nc * nbc[:, :, 1] + nc * nc * nbc[:, :, 2]
        cand_nb = members[cid].reshape(B, -1)                       # [B, 27*cap]
        valid = cand_nb >= 0
        bi, kk = np.nonzero(valid)
        ok = np.ones(B, dtype=bool)
        if bi.size:
            j = cand_nb[bi, kk]
            # a segment distance below dmin needs a centre distance below
            # (L_i + L_j)/2 + dmin: only those pairs run segdist2 (the
            # 1e-6 margin keeps every decision the full test would make)
            near = _centres_within(cr[bi], r[j], 0.5 * (cl[bi] + lengths[j]) + dmin + 1e-6, box)
            bi, j = bi[near], j[near]
        if bi.size:
            d2 = segdist2(cr[bi], cu[bi], cl[bi], r[j], u[j], lengths[j], box)
            bad = np.zeros(B, dtype=bool)
            np.logical_or.at(bad, bi[d2 < dmin * dmin], True)
            ok &= ~bad
        cand = np.nonzero(ok)[0][: N - placed]
        acc = []
        if cand.size:
            # conflicts among this batch's survivors, then a greedy pass
            ii, jj = np.triu_indices(cand.size, 1)
            conf = np.zeros((cand.size, cand.size), dtype=bool)
            if ii.size:
                near = _centres_within(cr[cand[ii]], cr[cand[jj]], 0.5 * (cl[cand[ii]] + cl[cand[jj]]) + dmin + 1e-6,
                                       box)
                ii, jj = ii[near], jj[near]
            if ii.size:
                a_, b_ = cand[ii], cand[jj]
                hit = segdist2(cr[a_], cu[a_], cl[a_], cr[b_], cu[b_], cl[b_], box) < dmin * dmin
                conf[ii[hit], jj[hit]] = True
                conf[jj[hit], ii[hit]] = True
            taken = np.zeros(cand.size, dtype=bool)
            for q in range(cand.size):
                if not conf[q, taken].any():
                    taken[q] = True
            acc = list(cand[taken])
        tried += B
        got += len(acc)
        if tried >= 20 * B and got < 0.02 * tried:
            # acceptance collapsed: narrow the axis distribution about +-z
            kappa = 2.0 if kappa == 0 else 2 * kappa
            tried = got = 0
        for b in acc:
            c = cc[b, 0] + nc * cc[b, 1] + nc * nc * cc[b, 2]
            if fill[c] == cap:
                members = np.concatenate([members, -np.ones((nc ** 3, cap), dtype=np.int64)], axis=1)
                cap *= 2
            r[placed], u[placed] = cr[b], cu[b]
            lengths[placed] = cl[b]
            members[c, fill[c]] = placed
            fill[c] += 1
            placed += 1
    if placed < N:
        raise RuntimeError("random sequential addition stalled at %d of %d rods" % (placed, N))
    return r, u


def place_native(seed, N, lengths, box, dmin=1.0):
    """Random sequential addition by the library's host routine
    msc_place_rods (its own generator, seeded by `seed`): the same rule as
    place_rsa, fast enough for a quarter of a million rods."""
    import ctypes
    from . import _lib
    lib = _lib.load()
    lengths = np.ascontiguousarray(lengths, dtype=np.float64)
    r = np.zeros((N, 3))
    u = np.zeros((N, 3))
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.msc_place_rods(N, box, lengths.ctypes.data_as(dp), dmin, seed, r.ctypes.data_as(dp),
                            u.ctypes.data_as(dp))
    if rc:
        raise RuntimeError("msc_place_rods could not place %d rods in a box of %.3f (code %d)" % (N, box, rc))
    return r, u


def place_aligned(rng, N, lengths, box, dmin=1.0, tilt=0.02):
    """Orthorhombic lattice of rods along +-z (random polarity, small random
    tilt): layers one longest rod apart, square in-plane grid."""
    c = lengths.max() + dmin + 0.05
    nz = max(1, int(box // c))
    nxy = int(math.ceil(math.sqrt(N / nz)))
    a = box / nxy
    if a < dmin * 1.02 or box / nz < c:
        raise RuntimeError("aligned lattice does not fit %d rods in a box of %.3f" % (N, box))
    idx = np.arange(nxy * nxy * nz)[:N]
    r = np.stack([(idx % nxy) * a, ((idx // nxy) % nxy) * a, (idx // (nxy * nxy)) * (box / nz)], axis=1)
    r += np.array([0.5 * a, 0.5 * a, 0.5 * box / nz]) - 0.5 * box
    u = fisher_axes(rng, N, 1.0 / (tilt * tilt))
    return r, u


def maxwell(rng, N, mass, inert, ori, temp):
    p = rng.normal(size=(N, 3)) * np.sqrt(mass * temp)[:, None]
    p -= p.mean(axis=0)
    lb = rng.normal(size=(N, 3)) * np.sqrt(inert * temp)
    l = np.einsum("nji,nj->ni", ori, lb)  # lab = O^T body
    return p, l


def inertia(length, dia, mass):
    """Spherocyl::set_inertia (v25/classes.cpp:175-185)."""
    R = dia / 2
    Rsq, Rft, Lsq = R * R, R ** 4, length * length
    vol = np.pi * Rsq * (length + 3.0 * R / 4)
    md = mass / vol
    I0 = md * np.pi * (Lsq * length * Rsq / 12 + Lsq * Rsq * R / 3 + length * Rft / 4 + 8.0 * Rft * R / 15)
    I2 = md * np.pi * (length * Rft / 2 + 8.0 * Rft * R / 15)
    return np.stack([I0, I0, I2], axis=1)


def morton_order(r, box, bits=10):
    """Permutation that numbers the centres r (in [0, box)^3) along a Morton
    (Z-order) curve over a 2^bits grid per axis: consecutive indices are
    neighbours in space, and each eighth of the index range is one octant
    (the pair kernels give each XCD one contiguous eighth of the particles)."""
    q = np.clip((np.mod(r, box) / box * (1 << bits)).astype(np.int64), 0, (1 << bits) - 1)
    code = np.zeros(len(r), dtype=np.int64)
    for b in range(bits - 1, -1, -1):
        for a in (2, 1, 0):  # z, y, x from the top: the first split is along z
            code = (code << 1) | ((q[:, a] >> b) & 1)
    return np.argsort(code, kind="stable")


def make_replica(rng, N, lengths, phi, mu, K, alpha_fn, rc, temp=1.0, dt=0.002, Q=5.0, gamma=0.0, seeds=(1, 2),
                 lattice=False, native_seed=None):
    dia = np.ones(N)
    mass = np.ones(N)
    vol = np.pi * dia ** 2 * lengths / 4 + np.pi * dia ** 3 / 6
    box = (vol.sum() / phi) ** (1.0 / 3)
    lengths = lengths.copy()
    if lattice:
        r, u = place_lattice(rng, N, lengths, box)
    elif native_seed is not None:
        r, u = place_native(native_seed, N, lengths, box)
    elif phi > 0.22:
        r, u = place_aligned(rng, N, lengths, box)
    else:
        r, u = place_rsa(rng, N, lengths, box)
    if not lattice and not phi > 0.22:
        # random addition numbers the rods in the order they were placed,
        # which is random in space: renumber them along a space-filling curve
        # (the same configuration, relabelled) so that neighbours in space
        # are neighbours in index
        o = morton_order(r, box)
        r, u, lengths = r[o], u[o], lengths[o]
    ori = frames(u)
    p, l = maxwell(rng, N, mass, inertia(lengths, dia, mass), ori, temp)
    dmax = float((lengths + dia).max())
    sr_cut = dmax + 0.5
    desc = replica_desc(N, box, alpha_fn(box), K, sr_cut, rc=rc, temp=temp, dt=dt, Q=Q, gamma=gamma,
                        dens=N / box ** 3, thermo_seed=seeds)
    state = dict(r=r, p=p, l=l, ori=ori, length=lengths, dia=dia, mass=mass, dpc=np.zeros((N, 1)),
                 doc=np.tile([0.0, 0.0, 1.0], (N, 1, 1)), dmu=np.full((N, 1), mu))
    return desc, state


def build(name):
    """[(ReplicaDesc, state dict), ...] of a workload."""
    if name == "c1":
        rng = np.random.default_rng(1)
        return [make_replica(rng, 256, np.full(256, 3.0), 0.10, math.sqrt(2.0), 6, lambda b: 5.6 / b, 0.0,
                             lattice=True)]
    if name == "c2":
        rng = np.random.default_rng(2)
        return [make_replica(rng, 4096, truncated_normal(rng, 4096, 3.0, 0.3), 0.20, 2.0, 8, lambda b: 7.0 / b, 0.0)]
    if name == "c3":
        rng = np.random.default_rng(3)
        return [make_replica(rng, 32768, truncated_normal(rng, 32768, 3.0, 0.3), 0.20, 2.0, 11,
                             lambda b: 3.0 / 3.5, 3.5)]
    if name == "c4":
        rng = np.random.default_rng(4)
        out = []
        for i in range(64):
            N = int(rng.integers(512, 8193))
            phi = float(rng.uniform(0.05, 0.30))
            mu = math.sqrt(float(rng.uniform(1.0, 6.0)))
            out.append(make_replica(rng, N, truncated_normal(rng, N, 3.0, 0.3), phi, mu, 8, lambda b: 3.0 / 3.5,
                                    3.5, seeds=(2 * i + 1, 2 * i + 2)))
        return out
    if name == "c5":
        rng = np.random.default_rng(5)
        return [make_replica(rng, 262144, truncated_normal(rng, 262144, 5.0, 0.5), 0.15, math.sqrt(6.0), 13,
                             lambda b: 3.0 / 3.5, 3.5, native_seed=5)]
    raise ValueError("unknown workload %r" % (name,))


def load_into(system, replicas):
    for k, (_, st) in enumerate(replicas):
        system.set_state(k, st["r"], st["p"], st["l"], st["ori"], st["length"], st["dia"], st["mass"],
                         st["dpc"], st["doc"], st["dmu"])
