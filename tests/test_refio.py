"""The reference's run files through the library: conf.dat read and written
by msc_read_conf / msc_write_conf (host-only, no GPU needed) and input.txt
read by refio.read_input, checked against the reference's own files of
v25/v251/run1 (tests/golden)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CONF = os.path.join(GOLD, "v251_run1_conf.txt")
INPUT = os.path.join(GOLD, "v251_run1_input.txt")


@pytest.fixture(scope="module")
def refio():
    import mdspherocyl_amd
    from mdspherocyl_amd import refio as r
    mdspherocyl_amd.load()
    return r


def test_conf_read_matches_tokens(refio):
    """Every value of the 500 particles, the box and both thermostats equals
    the number written in the file (parsed here independently)."""
    c = refio.read_conf(CONF, 500, 1)
    lines = open(CONF).read().splitlines()
    assert list(c["L"]) == [float(x) for x in lines[0].split()]
    for i in range(500):
        tok = lines[1 + i].split()
        bar = tok.index("|")
        vals = [float(x) for x in tok[:bar]]
        assert list(c["r"][i]) == vals[0:3] and list(c["p"][i]) == vals[3:6]
        assert list(c["ori"][i].ravel()) == vals[6:15] and list(c["l"][i]) == vals[15:18]
        assert c["mass"][i] == float(tok[bar + 1])
        assert list(c["dpos"][i, 0]) + list(c["dori"][i, 0]) == [float(x) for x in tok[bar + 2:bar + 8]]
        assert c["dmu"][i, 0] == float(tok[bar + 8]) and tok[bar + 9] == "."
        assert c["clust"][i] == int(tok[-1])
    th = [float(x) for x in lines[501].split() + lines[502].split()]
    assert list(c["thermo"]) == th


def test_conf_write_is_byte_identical(refio, tmp_path):
    """conf.dat read and written back reproduces the reference's file byte
    for byte (print_conf's formats: 6 significant digits on the box line,
    fixed 12 decimals after it)."""
    c = refio.read_conf(CONF, 500, 1)
    out = tmp_path / "conf.dat"
    refio.write_conf(out, c)
    assert out.read_bytes() == open(CONF, "rb").read()


def test_conf_rejects_a_wrong_count(refio):
    from mdspherocyl_amd import MscError
    with pytest.raises(MscError):
        refio.read_conf(CONF, 501, 1)
    with pytest.raises(MscError):
        refio.read_conf(CONF, 499, 1)
    with pytest.raises(MscError):
        refio.read_conf(CONF, 500, 2)


def test_read_input_v251(refio):
    """r_input's fields of v25/v251/run1/input.txt, and the box it implies
    equals the box line the same run wrote to conf.dat."""
    inp = refio.read_input(INPUT)
    assert (inp["Nsteps"], inp["pre_steps"], inp["signal"], inp["n_comp"]) == (101, 0, 0, 1)
    assert inp["n_part"] == [500] and inp["length"] == [9.0] and inp["diameter"] == [1.0] and inp["mass"] == [1.0]
    assert inp["B"] == [0.0, 0.0, 5.0]
    assert (inp["dens"], inp["temp"], inp["Q"], inp["gamma"]) == (0.0013, 1.0, 5.0, 1.0)
    assert (inp["ArLx"], inp["ArLy"], inp["Dt"], inp["sr_cut"]) == (1.0, 1.0, 1e-4, 10.6)
    assert (inp["K_max"], inp["alpha"], inp["cl_fr"], inp["min_s"], inp["iprint"]) == (5, 0.07, 10, 0.1, 10)
    assert inp["dip"] == [[(0.0, [0.0, 0.0, 1.0], 4.481)]]
    box = refio.box_from_input(inp)
    assert " ".join("%g" % x for x in box) == open(CONF).readline().strip()
    d = refio.replica_desc_from_input(inp)
    assert d.N == 500 and d.K == 5 and d.alpha == 0.07 and d.sr_cut == 10.6 and tuple(d.B) == (0.0, 0.0, 5.0)


def test_lab_dipoles_match_the_file(refio):
    """The lab-frame dipole columns of conf.dat equal what lab_dipoles forms
    from the particle's position and orientation with the input's body-frame
    dipole, to the 12 decimals written."""
    inp = refio.read_input(INPUT)
    c = refio.read_conf(CONF, 500, 1)
    _, _, dpc, doc, dmu = refio.species_arrays(inp)
    dpos, dori = refio.lab_dipoles(c["r"], c["ori"], dpc, doc)
    assert np.abs(dpos - c["dpos"]).max() <= 5e-13
    assert np.abs(dori - c["dori"]).max() <= 5e-13
    assert np.array_equal(dmu, c["dmu"])
    # an off-centre, tilted dipole: pos moves along the axis, ori rotates with the body
    rng = np.random.default_rng(3)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    r = rng.normal(size=(1, 3))
    dp, do = refio.lab_dipoles(r, q[None], np.array([[0.7]]), np.array([[[1.0, 0.0, 0.0]]]))
    assert np.allclose(dp[0, 0], r[0] + 0.7 * q[2]) and np.allclose(do[0, 0], q[0])


def test_species_arrays_two_species(refio):
    inp = {"n_part": [3, 2], "length": [2.0, 4.0], "diameter": [1.0, 0.5],
           "dip": [[(0.5, [0.0, 0.0, 1.0], 2.0)], [(-0.5, [1.0, 0.0, 0.0], 3.0)]]}
    length, dia, dpc, doc, dmu = refio.species_arrays(inp)
    assert list(length) == [2, 2, 2, 4, 4] and list(dia) == [1, 1, 1, 0.5, 0.5]
    assert list(dpc[:, 0]) == [0.5] * 3 + [-0.5] * 2 and list(dmu[:, 0]) == [2.0] * 3 + [3.0] * 2
    assert doc[4, 0].tolist() == [1.0, 0.0, 0.0]
    with pytest.raises(ValueError):
        refio.species_arrays(dict(inp, dip=[[(0.5, [0.0, 0.0, 1.0], 2.0)], []]))
