"""msc_pair_correl_noalloc (no device allocation, fixed device scratch)
against msc_pair_correl, through the C ABI.

Two calls in a row on one context, the second with a smaller input (fewer
particles, more species, fewer bars), so the second call runs on scratch
that still holds the first call's histograms.  Bar: the counts channel and
the dropped pairs equal exactly; the weighted channels (FP64 atomic sums in
an order that varies from run to run) within 1e-12 of the largest weight
times the count of their bar.  A request beyond the scratch limit fails
with MSC_E_CAPACITY before any kernel runs.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(msc, rng):
    from mdspherocyl_amd import replica_desc
    from oracle import oracle as O
    from tests import helpers as H
    sizes = [(2000, (16.0, 16.0, 16.0)), (600, (11.0, 12.0, 10.0))]
    sysm = msc.MdSystem([replica_desc(n, L, 0.5, 2, 5.0, rc=3.0) for n, L in sizes], precision="mixed")
    for rep, (n, L) in enumerate(sizes):
        S = O.OSys(n, 1, L=L)
        S.r[:] = rng.uniform(-1, 2, size=(n, 3)) * np.array(L)
        for i in range(n):
            q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
            S.ori[i] = q
        S.len[:] = 2.0
        S.set_inertia()
        H.load_osys(sysm, rep, S)
    return sysm, sizes


def _noalloc(sysm, rep, counts, bar_w, n_bars):
    from mdspherocyl_amd import _lib
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    nspec = len(counts)
    sets = nspec + nspec * (nspec - 1) // 2
    hist = np.zeros((sets, 4, n_bars))
    dropped = ctypes.c_long(-1)
    rc = sysm.L.msc_pair_correl_noalloc(sysm._ctx, rep, nspec, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                        bar_w, n_bars, hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        ctypes.byref(dropped))
    if rc != _lib.MSC_OK:
        raise _lib.MscError(rc, sysm.L.msc_last_error(sysm._ctx).decode())
    return hist, dropped.value


def test_pair_correl_noalloc_matches_allocating_entry_point():
    import mdspherocyl_amd as msc
    msc.load()
    rng = np.random.default_rng(23)
    sysm, sizes = _system(msc, rng)
    try:
        cases = [(0, [2000], 0.01), (1, [250, 350], 0.05)]   # the second input is smaller
        for rep, counts, bar_w in cases:
            from mdspherocyl_amd import refio
            nb = refio.correl_bars(sizes[rep][1], bar_w)
            print("case: noalloc pair correlations rep %d N %d species %d n_bars %d"
                  % (rep, sizes[rep][0], len(counts), nb), flush=True)
            hn, dn = _noalloc(sysm, rep, counts, bar_w, nb)
            ha, da = sysm.pair_correl(rep, bar_w, nb, counts=counts)
            assert dn == da
            assert np.array_equal(hn[:, 0], ha[:, 0])
            assert hn[:, 0].sum() > 0
            for c, wmax in ((1, 2.0), (2, 8.0), (3, 2.0)):
                assert np.all(np.abs(hn[:, c] - ha[:, c]) <= 1e-12 * wmax * np.maximum(ha[:, 0], 1))
        # 4 channels x 65537 bars of one species exceed the 262144-double scratch
        from mdspherocyl_amd import _lib
        with pytest.raises(_lib.MscError) as e:
            _noalloc(sysm, 1, [600], 1e-4, 65537)
        assert e.value.code == _lib.MSC_E_CAPACITY
    finally:
        sysm.close()
