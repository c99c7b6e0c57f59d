"""The reference-side adapter (mdspherocyl_amd/adapter) against the
reference's own declarations.

msc_v25_adapter.cpp must define each routine it replaces with exactly the
signature v25/functions.hpp declares, so that the reference's main.cpp links
against it unchanged.  Checked at the symbol level: a probe that takes the
address of every such routine, compiled from the declarations read out of the
reference's functions.hpp at test time, needs the same mangled symbols the
adapter's object file defines; and the reference's own definitions, renamed
with the -D options of the adapter's build recipe, define the ref_ symbols
the adapter calls.  The reference's headers need GSL (absent here), so both
compile over the test stand-ins of its containers (tests/adapter/mock_v25.hpp),
whose class names mangle as the reference's; every accessor the stand-ins
give the adapter is checked to exist in v25/classes.hpp with the same form.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/v25"
ADAPTER = os.path.join(ROOT, "mdspherocyl_amd", "adapter", "msc_v25_adapter.cpp")
MOCK_DIR = os.path.join(ROOT, "tests", "adapter")
ROUTINES = ["tabul", "wave_v", "neigh_list", "force_sc", "force_dipR", "force_dipK", "ext_field", "self_ener",
            "acc_correl", "oo_params"]
RENAMED = ["tabul", "wave_v", "force_dipR"]  # ref_ versions the adapter calls

needs_ref = pytest.mark.skipif(not os.path.exists(os.path.join(REF, "functions.hpp")),
                               reason="the reference's headers are not present")


def _declarations():
    """The declaration of each routine, as text, from v25/functions.hpp."""
    text = open(os.path.join(REF, "functions.hpp")).read()
    text = re.sub(r"//[^\n]*", "", text)
    out = {}
    for name in ROUTINES:
        m = re.search(r"(?:^|\n)\s*((?:bool|void|double)\s+%s\s*\([^;]*\)\s*;)" % name, text)
        assert m, name
        out[name] = m.group(1)
    return out


def _nm(obj, flag):
    res = subprocess.run(["nm", flag, "--format=posix", obj], capture_output=True, text=True, check=True)
    return {ln.split()[0]: ln.split()[1] for ln in res.stdout.splitlines() if ln.strip()}


def _compile(src, obj, inc):
    cmd = ["g++", "-std=c++17", "-c", src, "-o", obj] + ["-I" + d for d in inc]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


@needs_ref
def test_adapter_defines_the_reference_declarations(tmp_path):
    decls = _declarations()
    hdr = tmp_path / "functions.hpp"
    hdr.write_text("#pragma once\n#include \"mock_v25.hpp\"\n" + "\n".join(decls.values()) + "\n")
    inc = [str(tmp_path), MOCK_DIR, os.path.join(ROOT, "include"), os.path.join(ROOT, "mdspherocyl_amd", "adapter")]
    # what main.cpp needs: one undefined symbol per routine
    probe = tmp_path / "probe.cpp"
    probe.write_text("#include \"functions.hpp\"\nvoid *const msc_refs[] = {%s};\n"
                     % ", ".join("(void *)&%s" % n for n in ROUTINES))
    _compile(str(probe), str(tmp_path / "probe.o"), inc)
    need = {s for s, t in _nm(str(tmp_path / "probe.o"), "-u").items()}
    assert len(need) == len(ROUTINES), sorted(need)
    # what the adapter defines
    _compile(ADAPTER, str(tmp_path / "adapter.o"), inc)
    defined = {s for s, t in _nm(str(tmp_path / "adapter.o"), "--defined-only").items() if t == "T"}
    missing = need - defined
    assert not missing, "adapter lacks %s" % sorted(missing)
    # the renamed reference definitions supply every ref_ symbol the adapter calls
    ren = tmp_path / "renamed.cpp"
    ren.write_text("".join("#define %s ref_%s\n" % (n, n) for n in RENAMED) + "#include \"functions.hpp\"\n"
                   "void *const msc_refs[] = {%s};\n" % ", ".join("(void *)&ref_%s" % n for n in RENAMED))
    _compile(str(ren), str(tmp_path / "renamed.o"), inc)
    supplied = set(_nm(str(tmp_path / "renamed.o"), "-u"))
    called = {s for s in _nm(str(tmp_path / "adapter.o"), "-u") if "ref_" in s}
    assert called and called <= supplied, (sorted(called), sorted(supplied))


# accessors the adapter uses, as v25/classes.hpp and v25/Matrix.hpp declare them
ACCESSORS = [
    ("classes.hpp", r"int get_Ncomp\(\) const"),
    ("classes.hpp", r"int get_Tpart\(\) const"),
    ("classes.hpp", r"int get_Npart\(int comp\) const"),
    ("classes.hpp", r"double get_dip\(\) const"),
    ("classes.hpp", r"void get_pos\(double r\[3\]\) const"),
    ("classes.hpp", r"void get_ori\(double o\[3\]\) const"),
    ("classes.hpp", r"double get_mass\(\) const"),
    ("classes.hpp", r"void get_shape\(double sh\[2\]\) const"),
    ("classes.hpp", r"int get_ndips\(\) const"),
    ("classes.hpp", r"Dipole get_dip\(int i\) const"),
    ("classes.hpp", r"void get_mom\(double p\[3\]\) const"),
    ("classes.hpp", r"void get_angm\(double l\[3\]\) const"),
    ("classes.hpp", r"void get_omat\(Matrix<double>& o\) const"),
    ("classes.hpp", r"void get_lev\(int i, double d\[3\]\) const"),
    ("classes.hpp", r"void add_F\(const double f\[3\]\);"),
    ("classes.hpp", r"void add_T\(const double t\[3\]\);"),
    ("classes.hpp", r"Dynvec<int> C;"),
    ("classes.hpp", r"Dynvec<int> P;"),
    ("Matrix.hpp", r"void get_size\(int& rows, int& cols\) const"),
    ("Matrix.hpp", r"T const\* operator\[\]\(int i\) const"),
    ("Matrix.hpp", r"T& operator\[\]\(int i\)"),
    ("Matrix.hpp", r"Matrix\(int, int\);"),
]


@needs_ref
def test_stand_in_accessors_exist_in_the_reference():
    texts = {f: re.sub(r"[ \t]+", " ", open(os.path.join(REF, f)).read()) for f in ("classes.hpp", "Matrix.hpp")}
    for f, pat in ACCESSORS:
        assert re.search(pat, texts[f]), (f, pat)


def test_adapter_core_compiles_over_the_stand_ins(tmp_path):
    """The templated core alone (no reference headers needed), every routine
    instantiated."""
    src = tmp_path / "core.cpp"
    src.write_text(
        "#include \"mock_v25.hpp\"\n#include \"msc_adapter.hpp\"\n"
        "template class msc_adapter::Adapter<Dynvec<Dynvec<Spherocyl> >, part_num, Matrix<double> >;\n"
        "int use(msc_adapter::Adapter<Dynvec<Dynvec<Spherocyl> >, part_num, Matrix<double> > &a,\n"
        "        Dynvec<Dynvec<Spherocyl> > &p, part_num &np, Dynvec<Matrix<double> > &h, const cross *vc) {\n"
        "  double L[3] = {1, 1, 1};\n"
        "  return a.acc_correl(p, h, h, L, 0.1, np, vc);\n}\n")
    _compile(str(src), str(tmp_path / "core.o"),
             [MOCK_DIR, os.path.join(ROOT, "include"), os.path.join(ROOT, "mdspherocyl_amd", "adapter")])
