"""Distance files into per-position distance lists (cth_read_distances,
cth_add_distances), pinned by the reference's own transform: adddist's
ReadDistances and ProcessDistancesForPosition (src/adddist/adddist.cpp),
built from its source into oracle/_ref/adddist_ref by oracle/Makefile.ref.

Random distance files cover: several distance columns averaged, missing
values ('-'), probability columns with a threshold, comment and blank lines,
positions without lines, pairs inside the main-diagonal band, lists longer
than 99 entries (both truncation ends), and profile lengths around the
list limit.  The lists must agree value for value (bit-exact floats)."""
import os
import subprocess

import numpy as np
import pytest

import cother_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "adddist_ref")
OPTIONS = os.path.join(ROOT, "cother_amd", "env", "options.txt")

need_ref = pytest.mark.skipif(not os.path.exists(REF),
                              reason="reference adddist not built (oracle/Makefile.ref needs /root/reference)")


def _write_file(path, rng, L, ncol, with_prob, pmiss, band=1):
    """A distance file: for every i, pairs (i, j>i+band-1) in order, ncol
    distance columns (each with a probability when with_prob)."""
    lines = ["# distance file", "   # indented comment", ""]
    for i in range(1, L + 1):
        if rng.random() < 0.05:
            continue  # a position without lines
        for j in range(i + band, L + 1):
            if rng.random() < 0.3:
                continue
            cols = []
            for _ in range(ncol):
                if rng.random() < pmiss:
                    cols.append("-")
                    if with_prob:
                        cols.append(f"{rng.random():.3f}")
                else:
                    cols.append(f"{rng.uniform(2.0, 40.0):.2f}")
                    if with_prob:
                        cols.append(f"{rng.random():.3f}")
            sep = "\t" if rng.random() < 0.2 else " "
            lines.append(sep.join([str(i), str(j)] + cols))
            if rng.random() < 0.01:
                lines.append("# comment between lines")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _ref_lists(path, L, columns, prb):
    r = subprocess.run([REF, OPTIONS, path, str(L), ",".join(map(str, columns)), repr(prb)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    out = []
    for line in r.stdout.splitlines():
        t = line.split()
        out.append(np.array([float.fromhex(x) for x in t[1:]], np.float32))
        assert len(out[-1]) == int(t[0])
    assert len(out) == L
    return out


@need_ref
@pytest.mark.parametrize("seed,L,ncol,with_prob,prb,band", [
    (1, 60, 1, False, 0.0, 3),
    (2, 150, 1, False, 0.0, 1),
    (3, 130, 2, False, 0.0, 2),
    (4, 90, 1, True, 0.3, 3),
    (5, 120, 3, True, 0.5, 1),
    (6, 250, 1, False, 0.0, 3),
    (7, 8, 1, False, 0.0, 1),
])
def test_lists_match_reference(tmp_path, seed, L, ncol, with_prob, prb, band):
    rng = np.random.default_rng(seed)
    path = str(tmp_path / "d.txt")
    _write_file(path, rng, L, ncol, with_prob, 0.1, band)
    columns = [3 + k * (2 if with_prob else 1) for k in range(ncol)]
    want = _ref_lists(path, L, columns, prb)
    got = cother_amd.read_distances(path, L, columns, prb)
    assert len(got) == L
    for m in range(L):
        assert got[m].view(np.int32).tolist() == want[m].view(np.int32).tolist(), m
    assert max(len(x) for x in want) > 0


def test_add_distances_sets_records(tmp_path):
    ps = cother_amd.synth(2, 0, lmin=40, lmax=60, dist_mode=2, sort_desc=0, seed=3)
    L = int(ps.lens()[1])
    path = str(tmp_path / "d.txt")
    _write_file(path, np.random.default_rng(9), L, 1, False, 0.0, 3)
    lists = cother_amd.read_distances(path, L, [3], 0.0)
    cother_amd.add_distances(ps, 1, path, [3], 0.0)
    dd = ps.field("dd")
    d0 = int(ps.dists()[1])
    for m in range(L):
        rec = dd[d0 + m]
        assert rec[0] == len(lists[m])
        assert (rec[1:1 + rec[0]] == np.rint(lists[m]).astype(np.uint8)).all()
        assert (rec[1 + rec[0]:] == 255).all()
    # profile 0 is untouched
    assert (dd[:int(ps.lens()[0]), 0] == 0).all()


def test_invalid_files_are_refused(tmp_path):
    bad = {"decreasing": "5 9 3.0\n2 8 4.0\n", "range": "1 4 70.0\n", "beyond": "1 4 3.0\n2 99 3.0\n",
           "text": "1 4 x\n"}
    for name, text in bad.items():
        p = tmp_path / f"{name}.txt"
        p.write_text(text)
        with pytest.raises(RuntimeError):
            cother_amd.read_distances(str(p), 20, [3], 0.0)
