"""The C-ABI library loads on the CPU and exports every declared symbol."""
import ctypes
import os
import re

import cother_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(cth_\w+)\s*\(", src, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_library_exports_declared_symbols():
    L = cother_amd.lib()
    names = declared_functions()
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_declare_argument_types():
    """every declared entry point is bound with explicit argument types (no
    implicit int conversion of pointers)"""
    L = cother_amd.lib()
    unbound = [n for n in declared_functions() if getattr(L, n).argtypes is None]
    assert not unbound, unbound


def test_synth_deterministic():
    a = cother_amd.synth(5, 0, lmin=40, lmax=90, dist_mode=1, seed=3, threads=1)
    b = cother_amd.synth(5, 0, lmin=40, lmax=90, dist_mode=1, seed=3, threads=4)
    for f in ("trg", "trn", "ss", "dd", "bkg", "eno"):
        assert (a.field(f) == b.field(f)).all(), f
    assert (a.lens() == b.lens()).all()
    assert list(a.lens()) == sorted(a.lens(), reverse=True)


def test_synth_distance_lists():
    db = cother_amd.synth(20, 0, lmin=50, lmax=200, dist_mode=1, seed=4)
    dd = db.field("dd")
    n = dd[:, 0]
    assert n.max() <= 99
    # the last three positions of a profile never start a pair line -> empty
    ends = db.dists() + db.lens()
    for e in ends:
        assert (n[e - 3:e] == 0).all()
    # values beyond the count are the missing marker
    for i in range(0, db.npos, 97):
        assert (dd[i, 1 + n[i]:] == 255).all()


def test_concat_keeps_profiles():
    a = cother_amd.synth(2, 0, lmin=5, lmax=20, dist_mode=1, seed=1)
    b = cother_amd.synth(3, 0, lmin=5, lmax=20, dist_mode=0, seed=2)
    c = cother_amd.concat([a, b])
    assert list(c.lens()) == list(a.lens()) + list(b.lens())
    assert (c.field("trn")[:, a.npos + a.nprof:] == b.field("trn")).all()
    assert (c.field("dd")[a.npos:] == b.field("dd")).all()
    assert (c.field("bkg")[a.nprof:] == b.field("bkg")).all()
