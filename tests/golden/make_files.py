"""Writes the file fixtures of tests/test_gpu_files.py into tests/golden/files/.

- query.pro: one COTHER text profile (110 positions, no distance lines),
  written by this library's text writer (whose files the reference's makedb
  accepts, tests/test_proio.py);
- query.dst: a distance file in adddist's input format ("i j d" lines, pairs
  with j - i <= 60 and d <= 40 A) from a 3.8 A self-avoiding C-alpha chain;
- db.tbin.gz: a binary profile database written by the reference's own
  makedb and db2bin (built from its sources into oracle/_ref/ by
  oracle/Makefile.ref) from 14 text profiles: 12 random ones with C-alpha
  chain distance lists, the query with the lists adddist makes from
  query.dst (so the query has a true hit), and the query's first 70
  positions.

Run from the repository root in a container that has the reference build:
    python tests/golden/make_files.py
"""
import gzip
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import cother_amd  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "files")
REFDIR = os.path.join(ROOT, "oracle", "_ref")
OPTIONS = os.path.join(ROOT, "cother_amd", "env", "options.txt")


def chain(rng, n, step=3.8):
    """C-alpha coordinates of a self-avoiding random chain (rejection of steps
    closer than 3.8 A to any earlier atom)."""
    xyz = np.zeros((n, 3))
    for i in range(1, n):
        for _ in range(200):
            v = rng.normal(size=3)
            p = xyz[i - 1] + step * v / np.linalg.norm(v)
            if i < 3 or np.min(np.linalg.norm(xyz[:i - 1] - p, axis=1)) >= step:
                break
        xyz[i] = p
    return xyz


def write_dst(path, xyz, band=60):
    with open(path, "w") as f:
        f.write("# pairwise C-alpha distances (i j d), 1-based positions\n")
        n = len(xyz)
        for i in range(n):
            for j in range(i + 1, min(n, i + 1 + band)):
                d = np.linalg.norm(xyz[i] - xyz[j])
                if d <= 40.0:  # adddist takes distances in [0,64]
                    f.write(f"{i + 1} {j + 1} {d:.2f}\n")


def main():
    rng = np.random.default_rng(20261203)
    os.makedirs(OUT, exist_ok=True)
    q = cother_amd.synth(1, 2, lmin=110, lmax=110, dist_mode=2, sort_desc=0, seed=9101)
    cother_amd.set_text(q, 0, "query", "query profile of the file fixtures")
    cother_amd.write_text_profiles(q, os.path.join(OUT, "query.pro"))
    write_dst(os.path.join(OUT, "query.dst"), chain(rng, 110))
    tmp = tempfile.mkdtemp()
    try:
        files = []
        others = cother_amd.synth(12, 0, lmin=30, lmax=100, dist_mode=1, sort_desc=0, seed=9102)
        qd = cother_amd.read_text_profiles(os.path.join(OUT, "query.pro"))
        cother_amd.add_distances(qd, 0, os.path.join(OUT, "query.dst"))
        part = cother_amd.subset(qd, 0, 1)
        sets = [cother_amd.subset(others, k, k + 1) for k in range(others.nprof)] + [qd]
        # the query's first 70 positions as a profile of their own
        lens = np.array([70], np.int32)
        import ctypes
        head = cother_amd.Pset(cother_amd.lib().cth_pset_alloc(1, lens.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        for name in ("trg", "ss", "cv"):
            head.field(name)[:, :] = part.field(name)[:, :70]
        for name in ("cvprior", "cvnorm2", "aa", "sss", "dd"):
            head.field(name)[:] = part.field(name)[:70]
        head.field("trn")[:, :] = part.field("trn")[:, :71]
        head.field("bkg")[0] = part.field("bkg")[0]
        head.field("eno")[0] = part.field("eno")[0]
        sets.append(head)
        for k, s in enumerate(sets):
            cother_amd.set_text(s, 0, f"t{k}", f"template {k} of the file fixtures")
            f = os.path.join(tmp, f"t{k}.pro")
            cother_amd.write_text_profiles(s, f)
            files.append(f)
        db = os.path.join(tmp, "db")
        subprocess.run([os.path.join(REFDIR, "makedb_ref"), "-o", db, "-p", OPTIONS] + files, check=True, cwd=tmp,
                       capture_output=True)
        subprocess.run([os.path.join(REFDIR, "db2bin_ref"), "-d", db, "-p", OPTIONS], check=True, cwd=tmp,
                       capture_output=True)
        with open(db + ".tbin", "rb") as fi, gzip.GzipFile(os.path.join(OUT, "db.tbin.gz"), "wb", mtime=0) as fo:
            shutil.copyfileobj(fi, fo)
    finally:
        shutil.rmtree(tmp)
    for f in sorted(os.listdir(OUT)):
        print(f, os.path.getsize(os.path.join(OUT, f)))


if __name__ == "__main__":
    main()
