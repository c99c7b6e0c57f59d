"""Launch choices (msc_tuning) and the autotuner (msc_autotune) on the GPU.

System: two replicas of N=400 rods (L ~ N(3, 0.3), packing fraction 0.15,
mu^2 = 6, r_c = 3.5, alpha = 3/3.5), a small list skin so that list rebuilds
and contact-row refreshes happen within the runs, mixed precision (the only
precision with the fused pair launch).  Tolerances: choices documented as
exact are compared bit for bit; the waves-per-task choice changes the
summation order of the pair sums, so its runs agree to 1e-9 relative over
130 steps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def msc():
    import mdspherocyl_amd as m
    m.load()
    return m


def _system():
    from mdspherocyl_amd import configs
    rng = np.random.default_rng(41)
    N = 400
    d, st = configs.make_replica(rng, N, configs.truncated_normal(rng, N, 3.0, 0.3), 0.15, 1.5, 6,
                                 lambda b: 3.0 / 3.5, 3.5)
    d.dmin = 0.3
    return d, st


def _run(msc, tuning, autotune_steps=0):
    from mdspherocyl_amd import configs
    d, st = _system()
    sysm = msc.MdSystem([d, d], precision="mixed")
    try:
        configs.load_into(sysm, [(d, st), (d, st)])
        sysm.set_tuning(**tuning)
        sysm.init_run()
        tuned = sysm.autotune(autotune_steps) if autotune_steps else None
        o1 = sysm.run(70, step0=40)
        o2 = sysm.run(60, step0=110)
        F, T = sysm.get_forces(0)
        return np.concatenate([o1, o2]), sysm.get_state(0), (F, T), tuned, sysm.get_tuning()
    finally:
        sysm.close()


def test_fused_pairs_equal_separate_launches(msc):
    """The fused WCA + real-space launch forms every sum as the two separate
    launches do: 130 steps, observables, state and forces bit for bit."""
    _say("case: fused vs separate pair launches, mixed, 2 x N=400, 130 steps")
    a = _run(msc, dict(fused_pairs=1, pair_waves=4))
    b = _run(msc, dict(fused_pairs=0))
    assert int(a[0][:, 0, 11].sum()) >= 1
    assert np.array_equal(a[0], b[0])
    for x, y in zip(a[1], b[1]):
        assert np.array_equal(x, y)
    for x, y in zip(a[2], b[2]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("graph", [0, 1])
def test_overlapped_pair_stream_equals_one_stream(msc, graph):
    """The pair launch on a second stream, concurrent with the structure
    factors and their reduction, gives the one-stream step bit for bit
    (direct launches and the captured graph): 130 steps, observables, state
    and forces."""
    _say("case: overlapped vs one stream, graph %d, mixed, 2 x N=400, 130 steps" % graph)
    a = _run(msc, dict(fused_pairs=1, pair_waves=4, graph=graph, overlap=1))
    b = _run(msc, dict(fused_pairs=1, pair_waves=4, graph=graph, overlap=0))
    assert a[4]["overlap"] == 1 and b[4]["overlap"] == 0
    assert np.array_equal(a[0], b[0])
    for x, y in zip(a[1], b[1]):
        assert np.array_equal(x, y)
    for x, y in zip(a[2], b[2]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("graph", [0, 1])
def test_fused_last_kick_equals_kick_launch(msc, graph):
    """The last kick in the back projection's epilogue (thermostat
    coordinates from the reduction, cell fill in its extra blocks) gives the
    step with the kick as a launch of its own bit for bit (direct launches
    and the captured graph): 130 steps with list rebuilds, observables,
    state and forces."""
    _say("case: fused last kick vs kick launch, graph %d, mixed, 2 x N=400, 130 steps" % graph)
    a = _run(msc, dict(fused_pairs=1, pair_waves=4, graph=graph, fuse_kick=1))
    b = _run(msc, dict(fused_pairs=1, pair_waves=4, graph=graph, fuse_kick=0))
    assert a[4]["fuse_kick"] == 1 and b[4]["fuse_kick"] == 0
    assert int(a[0][:, 0, 11].sum()) >= 1
    assert np.array_equal(a[0], b[0])
    for x, y in zip(a[1], b[1]):
        assert np.array_equal(x, y)
    for x, y in zip(a[2], b[2]):
        assert np.array_equal(x, y)


def test_own_rows_follow_reference_rows(msc):
    """The context's own rows (list skin 1.0 and 0.5) against the
    reference's rows and rebuild test (list skin 0): the same pairs interact,
    in another summation order, so the trajectories agree to 1e-9 over 130
    steps; both schedules rebuild within the run."""
    _say("case: own rows vs reference rows, mixed, 2 x N=400, 130 steps")
    a = _run(msc, dict(fused_pairs=1, pair_waves=4, list_skin=0.0))
    for skin in (1.0, 0.5):
        b = _run(msc, dict(fused_pairs=1, pair_waves=4, list_skin=skin))
        assert b[4]["list_skin"] == skin
        _say("  skin %.2f: rebuilds %d (reference rows: %d)"
             % (skin, int(b[0][:, 0, 11].sum()), int(a[0][:, 0, 11].sum())))
        assert int(b[0][:, 0, 11].sum()) >= 1 and int(a[0][:, 0, 11].sum()) >= 1
        assert np.allclose(b[0][:, :, 0], a[0][:, :, 0], rtol=1e-9, atol=0)
        assert np.allclose(b[0][:, :, 1], a[0][:, :, 1], rtol=1e-9, atol=1e-12)
        F, Fb = a[2][0], b[2][0]
        assert np.abs(F - Fb).max() <= 1e-9 * np.abs(F).max()


@pytest.mark.parametrize("chunk", [64, 512])
def test_chunk_changes_rounding_only(msc, chunk):
    """Structure-factor chunks of 64 or 512 particles against 128.  The
    chunk partials of a mixed context are FP32 sums (added in FP64), so
    another grouping changes the reciprocal energy and forces by FP32
    rounding: over 130 steps K agrees to 1e-9 relative, U per particle (about
    0.1, a small sum of larger parts) to 1e-7, forces to 1e-6 of the largest."""
    _say("case: chunk %d vs 128, mixed, 2 x N=400, 130 steps" % chunk)
    a = _run(msc, dict(fused_pairs=1, pair_waves=4, chunk=128))
    b = _run(msc, dict(fused_pairs=1, pair_waves=4, chunk=chunk))
    assert b[4]["chunk"] == chunk
    dK = np.abs(b[0][:, :, 0] / a[0][:, :, 0] - 1).max()
    dU = np.abs(b[0][:, :, 1] - a[0][:, :, 1]).max()
    F, Fb = a[2][0], b[2][0]
    dF = np.abs(F - Fb).max() / np.abs(F).max()
    _say("  K %.2e relative, U %.2e absolute, F %.2e of the largest" % (dK, dU, dF))
    assert dK <= 1e-9 and dU <= 1e-7 and dF <= 1e-6


@pytest.mark.parametrize("waves", [2, 8])
def test_pair_waves_change_rounding_only(msc, waves):
    """2 or 8 waves per task against 4: the same trajectory to 1e-9."""
    _say("case: %d waves per task vs 4, mixed, 2 x N=400, 130 steps" % waves)
    a = _run(msc, dict(fused_pairs=1, pair_waves=4))
    b = _run(msc, dict(fused_pairs=1, pair_waves=waves))
    K, U = a[0][:, :, 0], a[0][:, :, 1]
    assert np.allclose(b[0][:, :, 0], K, rtol=1e-9, atol=0)
    assert np.allclose(b[0][:, :, 1], U, rtol=1e-9, atol=1e-12)
    F, Fb = a[2][0], b[2][0]
    assert np.abs(F - Fb).max() <= 1e-9 * np.abs(F).max()


def test_autotune_keeps_the_run(msc):
    """msc_autotune tries every candidate from the state at the call and
    leaves that state (lists rebuilt from it): the run afterwards follows the
    untuned run (K to 1e-9 relative, U to 1e-7 per particle); the choice is
    one of the candidates."""
    _say("case: autotune over 20 steps, then 130 steps, mixed, 2 x N=400")
    a = _run(msc, dict(fused_pairs=1, pair_waves=4))
    b = _run(msc, dict(fused_pairs=1, pair_waves=4), autotune_steps=20)
    best, tried = b[3]
    _say("  tried: %s" % ", ".join("%s %.4f ms" % (c, t) for c, t in tried))
    _say("  chose: %s" % best)
    # 8 launch choices, the other overlap choice, the three list skins and
    # three chunks other than the defaults
    assert len(tried) == 15 and all(t > 0 for _, t in tried)
    assert sorted({c["overlap"] for c, _ in tried}) == [0, 1]
    assert sorted({c["list_skin"] for c, _ in tried}) == [0.5, 0.75, 1.0, 1.5]
    assert sorted({c["chunk"] for c, _ in tried}) == [64, 128, 256, 512]
    assert best in [c for c, _ in tried]
    assert b[4] == best
    # a chosen chunk size regroups the FP32 chunk partials (see
    # test_chunk_changes_rounding_only): U to 1e-7 per particle
    assert np.allclose(b[0][:, :, 0], a[0][:, :, 0], rtol=1e-9, atol=0)
    assert np.abs(b[0][:, :, 1] - a[0][:, :, 1]).max() <= 1e-7
