"""The results files, plain text and JSON (cth_format_hit_*, cth_format_results_*,
cth_write_results_*), pinned by the reference's own results writer.

- whole files: the reference's AlnWriter (cualn/AlnWriter.cpp,
  AlnWriterPlain.cpp), built from its sources into oracle/_ref/ by
  oracle/Makefile.ref, receives our per-hit records split into parts, merges
  them and writes the file (plain, or JSON with B_FMT set); our writer must
  produce the same bytes, for empty and non-empty hit lists, NOHITS/NOALNS
  limits, long database names and several parts;
- descriptions: the annotation and record headers equal the reference's
  FormatDescription (liblib/fmtdescription.h) applied to the same text;
- records: the formatting of the finalizer (cupro/CuBatchProcessingFinalizerPlain.cpp,
  which includes CUDA headers and so cannot be built here) is checked by its
  invariants: the alignment blocks re-assemble to the hit's rows, the line
  coordinates follow from the begin coordinates and the gaps, and the counts
  and percentages are those of the hit.  The exact printf formats of the
  score and footer lines are transcribed (parity unpinned beyond the
  invariants).
"""
import math
import os
import re
import subprocess

import numpy as np
import pytest

import cother_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFDIR = os.path.join(ROOT, "oracle", "_ref")
WRITER = os.path.join(REFDIR, "alnwriter_ref")
OPTIONS = os.path.join(ROOT, "cother_amd", "env", "options.txt")
PARFILE = os.path.join(ROOT, "cother_amd", "env", "cother.par")

need_ref = pytest.mark.skipif(not os.path.exists(WRITER),
                              reason="reference writer not built (oracle/Makefile.ref needs /root/reference)")

WORDS = ["alpha", "beta-sheet", "helix", "domain", "protein", "kinase", "(fragment)", "Homo", "sapiens",
         "transferase", "X" * 30, "putative", "uncharacterized", "2.7.11.1", "ATP-binding", "N-terminal"]


def _desc(rng, nwords):
    sep = ["  ", " ", "\t", " "]
    out = []
    for k in range(nwords):
        out.append(WORDS[rng.integers(len(WORDS))])
        out.append(sep[rng.integers(len(sep))])
    return "".join(out).strip()


def _rows(rng, alnlen, qss, tss):
    """Five alignment rows: query, middle, target, query SS, target SS."""
    aa = np.frombuffer(b"ARNDCQEGHILKMFPSTWYV", np.uint8)
    q = aa[rng.integers(20, size=alnlen)].copy()
    t = aa[rng.integers(20, size=alnlen)].copy()
    # gaps: never in both rows at one column; the first and last columns are matches
    g = rng.random(alnlen)
    q[(g < 0.08)] = ord("-")
    t[(g > 0.92)] = ord("-")
    q[0] = t[0] = q[-1] = t[-1] = ord("A")
    mid = np.full(alnlen, ord(" "), np.uint8)
    same = q == t
    mid[same] = q[same]
    plus = (~same) & (q != ord("-")) & (t != ord("-")) & (rng.random(alnlen) < 0.3)
    mid[plus] = ord("+")
    ssl = np.frombuffer(b" eh", np.uint8)
    qs = ssl[rng.integers(3, size=alnlen)] if qss else np.full(alnlen, ord(" "), np.uint8)
    ts = ssl[rng.integers(3, size=alnlen)] if tss else np.full(alnlen, ord(" "), np.uint8)
    return [bytes(r) for r in (q, mid, t, qs, ts)]


def _hits(rng, n, db, lq):
    hits = np.zeros(n, cother_amd.HIT)
    rows = []
    lens = db.lens()
    for k in range(n):
        p = int(rng.integers(db.nprof))
        alnlen = int(rng.integers(1, 200))
        r = _rows(rng, alnlen, True, True)
        rows.append(r)
        h = hits[k]
        h["query"] = 0
        h["profile"] = p
        h["score"] = rng.uniform(5, 300)
        h["logE"] = rng.uniform(-60, 2.3)
        h["alnlen"] = alnlen
        h["idnts"] = sum(1 for a, b in zip(r[0], r[2]) if a == b and a != ord("-"))
        h["pstvs_map"] = h["idnts"] + r[1].count(b"+")
        h["ngaps"] = r[0].count(b"-") + r[2].count(b"-")
        h["mapbeg"] = (int(rng.integers(0, max(1, lens[p] - 1))) << 16) | int(rng.integers(0, max(1, lq - 1)))
        h["expected"] = rng.choice([-0.5, 0.3])
        h["K"], h["lambda_"] = rng.choice([0.1, -1.0]), rng.choice([0.3, -1.0])
        h["K_est"], h["lambda_est"] = rng.uniform(0.01, 0.2), rng.uniform(0.1, 0.6)
        h["H"], h["minsc"], h["maxsc"] = rng.uniform(0, 2), -rng.uniform(1, 20), rng.uniform(1, 20)
    # distinct log E-values: the reference's unstable sort leaves ties unordered
    hits["logE"] = np.round(hits["logE"], 3) + np.arange(n) * 1e-4
    return hits, rows


def _alnbuf(hits, rows):
    buf = bytearray()
    for h, r in zip(hits, rows):
        h["alnoff"] = len(buf)
        for row in r:
            buf += row
    return bytes(buf)


@pytest.fixture(scope="module")
def sets():
    qs = cother_amd.synth(2, 0, lmin=60, lmax=300, dist_mode=2, sort_desc=0, seed=41)
    db = cother_amd.synth(40, 0, lmin=30, lmax=700, dist_mode=2, sort_desc=1, seed=42)
    rng = np.random.default_rng(7)
    for p in range(db.nprof):
        cother_amd.set_text(db, p, f"d{p}", _desc(rng, int(rng.integers(0, 60))))
    cother_amd.set_text(qs, 0, "myquery.pro", _desc(rng, 25))
    cother_amd.set_text(qs, 1, "other.query.tpro", "")
    return qs, db


def _ref_file(tmp_path, qs, db, hits, rows, info, nparts, options=OPTIONS, fmt="plain"):
    """The reference writer's file for our per-hit records in nparts parts."""
    lq = int(qs.lens()[0])
    ss, dl = cother_amd.search_space(lq, info["dbsize"], info["ndbentries"])
    recs = [cother_amd.format_hit_plain(qs, 0, db, h, r, fmt=fmt, **info) for h, r in zip(hits, rows)]
    # the finalizer hands over only hits below the threshold (CompressResultsPlain)
    order = np.array([k for k in range(len(hits)) if not (np.float32(info["logevthld"]) < hits[k]["logE"])], int)
    parts = np.array_split(order, nparts) if len(order) else [order]
    inp = bytearray()
    inp += f"{lq} {dl} {float(np.float32(ss)).hex()} {float(np.float32(info['logevthld'])).hex()}\n".encode()
    inp += (cother_amd.pset_name(qs, 0) or "query").encode() + b"\n"
    inp += (cother_amd.pset_desc(qs, 0) or "").encode() + b"\n"
    inp += f"{info['dbsize']} {info['ndbentries']} {len(info['dbnames'])}\n".encode()
    for n in info["dbnames"]:
        inp += n.encode() + b"\n"
    inp += f"{len(parts)}\n".encode()
    for part in parts:
        inp += f"{len(part)}\n".encode()
        for k in part:
            a, b = recs[k]
            inp += f"{float(hits[k]['logE']).hex()} {len(a)} {len(b)}\n".encode() + a + b
    d = tmp_path / f"ref{nparts}{fmt}"
    d.mkdir(exist_ok=True)
    f = tmp_path / f"in{nparts}{fmt}.txt"
    f.write_bytes(bytes(inp))
    flag = ["--json"] if fmt == "json" else []
    r = subprocess.run([WRITER, *flag, options, PARFILE, str(d), str(f)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    files = os.listdir(d)
    assert len(files) == 1, files
    return files[0], (d / files[0]).read_bytes()


@need_ref
@pytest.mark.parametrize("fmt", ["plain", "json"])
@pytest.mark.parametrize("nhits,nparts", [(0, 1), (1, 1), (7, 2), (25, 3)])
def test_file_matches_reference_writer(tmp_path, sets, nhits, nparts, fmt):
    qs, db = sets
    rng = np.random.default_rng(100 + nhits)
    hits, rows = _hits(rng, nhits, db, int(qs.lens()[0]))
    buf = _alnbuf(hits, rows)
    info = dict(dbnames=["/data/dbs/pdb70_2021", "scop"], dbsize=int(db.npos), ndbentries=db.nprof,
                logevthld=math.log(10.0))
    name, want = _ref_file(tmp_path, qs, db, hits, rows, info, nparts, fmt=fmt)
    ours = tmp_path / "ours"
    ours.mkdir()
    path = cother_amd.write_results_plain(qs, 0, db, hits, buf, str(ours), fmt=fmt, **info)
    assert os.path.basename(path) == name
    got = open(path, "rb").read()
    if fmt == "json":
        import json
        doc = json.loads(got)["cother_search"]  # valid JSON when NOALNS >= NOHITS
        assert len(doc["search_hits"]) == len(doc["search_summary"])
    if got != want:
        a, b = got.decode(errors="replace").splitlines(), want.decode(errors="replace").splitlines()
        for i, (x, y) in enumerate(zip(a, b)):
            assert x == y, f"line {i}: ours {x!r} reference {y!r}"
        assert len(a) == len(b), f"{len(a)} lines vs {len(b)}"
        assert got == want


@need_ref
@pytest.mark.parametrize("fmt", ["plain", "json"])
def test_limits_and_long_names(tmp_path, sets, fmt):
    """NOHITS/NOALNS below the number of hits, a long database name list."""
    qs, db = sets
    rng = np.random.default_rng(5)
    hits, rows = _hits(rng, 12, db, int(qs.lens()[0]))
    buf = _alnbuf(hits, rows)
    opts = open(OPTIONS).read()
    opts = re.sub(r"(?m)^NOHITS = \d+", "NOHITS = 5", opts)
    opts = re.sub(r"(?m)^NOALNS = \d+", "NOALNS = 3", opts)
    opts = re.sub(r"(?m)^ALNWIDTH = \d+", "ALNWIDTH = 37", opts)
    opts = re.sub(r"(?m)^DSCWIDTH = \d+", "DSCWIDTH = 50", opts)
    of = tmp_path / "options.txt"
    of.write_text(opts)
    names = ["/x/" + "n" * 130] + [f"db_{k}_" + "q" * 40 for k in range(30)]
    info = dict(dbnames=names, dbsize=int(db.npos), ndbentries=db.nprof, logevthld=math.log(0.5),
                nohits=5, noalns=3, alnwidth=37, dscwidth=50)
    _, want = _ref_file(tmp_path, qs, db, hits, rows, info, 2, options=str(of), fmt=fmt)
    got = cother_amd.format_results_plain(qs, 0, db, hits, buf, fmt=fmt, **info)
    assert got == want


@need_ref
def test_descriptions_match_reference_wrapping(sets):
    qs, db = sets
    rng = np.random.default_rng(9)
    for p in range(db.nprof):
        hits, rows = _hits(rng, 1, db, int(qs.lens()[0]))
        hits[0]["profile"] = p
        for dscwidth in (80, 33):
            ann, rec = cother_amd.format_hit_plain(qs, 0, db, hits[0], rows[0], dscwidth=dscwidth)
            desc = cother_amd.pset_desc(db, p).encode()
            ref_ann = subprocess.run([WRITER, "--fmt", "75", "0", "75", "0"], input=desc, capture_output=True).stdout
            ref_rec = subprocess.run([WRITER, "--fmt", "1024", "13", str(dscwidth), str(ord(">"))], input=desc,
                                     capture_output=True).stdout
            assert ann[:len(ref_ann)] == ref_ann
            assert len(ann) == max(75, len(ref_ann)) + 17
            assert rec.startswith(ref_rec + b"\n  Length/ENO: ")


def test_record_invariants(sets):
    qs, db = sets
    rng = np.random.default_rng(11)
    hits, rows = _hits(rng, 20, db, int(qs.lens()[0]))
    for h, r in zip(hits, rows):
        _, rec = cother_amd.format_hit_plain(qs, 0, db, h, r, alnwidth=60)
        text = rec.decode()
        n = int(h["alnlen"])
        m = re.search(r"Identities = (\d+)/(\d+) \((\d+)%\)", text)
        if h["idnts"] > 0:
            assert (int(m[1]), int(m[2]), int(m[3])) == (h["idnts"], n, h["idnts"] * 100 // n)
        else:
            assert m is None
        m = re.search(r"Gaps = (\d+)/(\d+) \((\d+)%\)", text)
        assert (m is not None) == (h["ngaps"] > 0)
        e = math.exp(float(h["logE"]))
        assert f"Expect = {e:.2g}," in text
        # blocks re-assemble to the rows; coordinates advance by the non-gap letters
        qline = re.findall(r"(?m)^Query: +(\d+) (\S+) (\d+) *$", text)
        tline = re.findall(r"(?m)^Sbjct: +(\d+) (\S+) (\d+) *$", text)
        assert "".join(s for _, s, _ in qline).encode() == r[0]
        assert "".join(s for _, s, _ in tline).encode() == r[2]
        assert all(len(s) <= 60 for _, s, _ in qline)
        for lines, beg in ((qline, (int(h["mapbeg"]) & 0xffff) + 1), (tline, (int(h["mapbeg"]) >> 16) + 1)):
            pos = beg
            for b, s, e_ in lines:
                assert int(b) == pos
                k = len(s) - s.count("-")
                pos += k
                assert int(e_) == (pos if k == 0 else pos - 1)
        lines = text.split("\n")
        mids = [lines[i + 1] for i, x in enumerate(lines) if x.startswith("Query:")]
        assert all(x.startswith(" " * 13) for x in mids)
        assert "".join(x[13:] for x in mids).encode() == r[1]
        ss = re.findall(r"(?m)^struct {7}(.*)$", text)
        assert "".join(ss[0::2]).encode() == r[3] and "".join(ss[1::2]).encode() == r[4]


def test_no_ss_rows_without_predictions(sets):
    qs, db = sets
    nss = cother_amd.synth(3, 0, lmin=40, lmax=60, dist_mode=2, sort_desc=0, seed=3)
    nss.field("sss")[:] = 0
    rng = np.random.default_rng(2)
    hits, rows = _hits(rng, 1, nss, int(qs.lens()[0]))
    _, rec = cother_amd.format_hit_plain(qs, 0, nss, hits[0], rows[0])
    assert b"struct" not in rec
    hits, rows = _hits(rng, 1, db, int(qs.lens()[0]))
    _, rec = cother_amd.format_hit_plain(qs, 0, db, hits[0], rows[0])
    assert b"struct" in rec
    _, rec = cother_amd.format_hit_plain(qs, 0, db, hits[0], rows[0], ssswgt=0.0)
    assert b"struct" not in rec


def test_search_space_matches_oracle():
    from tests import _oracle
    o = _oracle.load()
    for m, n, N in ((150, 450000, 2000), (40, 14_000_000, 60000), (1000, 300, 3), (5, 10, 1)):
        assert cother_amd.search_space(m, n, N) == o.length_adjustment(m, n, N)
