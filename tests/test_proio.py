"""The reference's profile file formats: the text profile writer and reader and
the binary database reader, pinned by the reference's own database tools
(makedb, db2bin) built from its sources into oracle/_ref/ by
oracle/Makefile.ref.

- our text writer's profiles are accepted by the reference's makedb, whose
  text database (written by the reference's TextWriteProfileCondensed) reads
  back to the same profile set;
- the reference's db2bin turns that text database into the binary format,
  which our binary reader reads to exactly the values our text reader gives
  for the text database: the two readers agree with the reference's own
  parse and binary writer, field by field and bit for bit.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import cother_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFDIR = os.path.join(ROOT, "oracle", "_ref")
MAKEDB = os.path.join(REFDIR, "makedb_ref")
DB2BIN = os.path.join(REFDIR, "db2bin_ref")
OPTIONS = os.path.join(ROOT, "cother_amd", "env", "options.txt")

FIELDS = ("trg", "ss", "cv", "cvprior", "cvnorm2", "trn", "bkg", "eno", "dd", "aa", "sss")


def _same(a, b, skip=()):
    assert list(a.lens()) == list(b.lens())
    for f in FIELDS:
        if f in skip:
            continue
        x, y = a.field(f), b.field(f)
        if x.dtype.kind == "f":
            bad = np.argwhere(x.view(np.int32) != y.view(np.int32))
        else:
            bad = np.argwhere(x != y)
        assert bad.size == 0, f"{f}: {len(bad)} values differ, first at {bad[0].tolist()}: {x[tuple(bad[0])]} vs {y[tuple(bad[0])]}"


@pytest.fixture(scope="module")
def sets():
    from tests._cv import set_informative_cv
    # the third set carries context vectors ("CV:" lines) at about half its positions
    return [cother_amd.synth(6, 0, lmin=5, lmax=120, dist_mode=1, sort_desc=0, seed=201),
            cother_amd.synth(4, 0, lmin=40, lmax=90, dist_mode=0, sort_desc=0, seed=202),
            set_informative_cv(cother_amd.synth(3, 0, lmin=20, lmax=70, dist_mode=1, sort_desc=0, seed=203), 204,
                               frac=0.5)]


def _ss_written(v, scale=10000):
    """SS probabilities after the writer's floor((float)scale * v) and the reader's
    (float)k / (float)scale (PMProfileModel.cpp:2046-2056, 4470-4482)."""
    f = np.float32(scale)
    return (np.trunc(f * v.astype(np.float32)).astype(np.int64).astype(np.float32) / f).astype(np.float32)


def test_text_round_trip(tmp_path, sets):
    for k, ps in enumerate(sets):
        path = str(tmp_path / f"s{k}.pro")
        cother_amd.write_text_profiles(ps, path)
        back = cother_amd.read_text_profiles(path)
        _same(ps, back, skip=("ss",))
        for p in range(ps.nprof):
            assert cother_amd.pset_name(back, p) == cother_amd.pset_name(ps, p)
            assert cother_amd.pset_desc(back, p) == cother_amd.pset_desc(ps, p)
        # the reference writes SS probabilities with a floor: a value may drop by 1/scale
        assert (back.field("ss").view(np.int32) == _ss_written(ps.field("ss")).view(np.int32)).all()


@pytest.mark.skipif(not (os.path.exists(MAKEDB) and os.path.exists(DB2BIN)),
                    reason="reference database tools not built (oracle/Makefile.ref needs /root/reference)")
def test_reference_tools_agree(tmp_path, sets):
    ps = cother_amd.concat(sets)
    files = []
    for p in range(ps.nprof):
        one = cother_amd.subset(ps, p, p + 1)
        f = str(tmp_path / f"p{p}.pro")
        cother_amd.write_text_profiles(one, f)
        files.append(f)
    db = str(tmp_path / "db")
    r = subprocess.run([MAKEDB, "-o", db, "-p", OPTIONS] + files, capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([DB2BIN, "-d", db, "-p", OPTIONS], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    text = cother_amd.read_text_profiles(db + ".tprd")
    binary = cother_amd.read_binary_db(db + ".tbin")
    # both readers see the reference's own parse and binary writer identically
    _same(text, binary)
    # makedb keeps profiles in its own order; match them by content.  The
    # reference tools rewrite SS probabilities twice with a floor (.pro ->
    # .tprd -> binary), everything else is carried exactly
    order = _match(ps, text)
    src = cother_amd.subset_list(ps, order)
    _same(src, text, skip=("ss",))
    # descriptions travel through makedb (text database) and db2bin (description section)
    for p in range(src.nprof):
        assert cother_amd.pset_desc(text, p) == cother_amd.pset_desc(src, p)
        assert cother_amd.pset_desc(binary, p) == cother_amd.pset_desc(src, p)
    assert (text.field("ss").view(np.int32) == _ss_written(_ss_written(src.field("ss"))).view(np.int32)).all()


def _match(a, b):
    """Index in a of each profile of b (by length and residues)."""
    la, lb = a.lens(), b.lens()
    da, db_ = a.dists(), b.dists()
    aa_a, aa_b = a.field("aa"), b.field("aa")
    out = []
    for j in range(b.nprof):
        cand = [i for i in range(a.nprof) if la[i] == lb[j] and i not in out and
                (aa_a[da[i]:da[i] + la[i]] == aa_b[db_[j]:db_[j] + lb[j]]).all()]
        assert cand, f"profile {j} of the database has no source"
        out.append(cand[0])
    return out
