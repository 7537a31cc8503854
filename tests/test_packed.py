"""The ingest of the reference's packed profile-model buffers
(cth_pset_from_packed, cth_upload_db_packed, cth_upload_queries_packed),
the form in which the reference's reader hands a database chunk to
CuBatchProcessing::TransferCPMData (libmycu/cupro/CuBatchProcessing.cuh:
62-70) and its workers hand a query to ProcessScoreMatrix.

- a profile set packed field by field as TdDataReader::ReadDataChunkHelper
  packs a chunk (32-byte aligned fields in one buffer, offsets in
  szCpm2dvfields) reads back to the same values;
- the sections of a db2bin file written by the reference's own tools, copied
  into a chunk buffer as the reference's reader copies them (a chunk from the
  middle of the database), read back to what the binary reader gives for
  those profiles;
- a query read from the middle of a packed batch (pointers advanced past the
  profiles before it, as PMDataSkipNPros advances them);
- malformed buffers are refused with an error.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import cother_amd
from tests import _packed
from tests.test_proio import DB2BIN, MAKEDB, OPTIONS, _same


def from_packed(pk, nprof=-1, szfields=True, advance=None):
    L = cother_amd.lib()
    beg, end = pk.beg, pk.end
    if advance is not None:
        beg = (ctypes.c_void_p * _packed.NFIELDS)(*advance)
    p = L.cth_pset_from_packed(beg, end if nprof < 0 else None, pk.sz if szfields else None, nprof, None)
    if not p:
        raise RuntimeError(L.cth_last_error().decode())
    return cother_amd.Pset(p)


@pytest.fixture(scope="module")
def sets():
    from tests._cv import set_informative_cv
    a = cother_amd.synth(7, 0, lmin=5, lmax=120, dist_mode=1, sort_desc=0, seed=301)
    b = cother_amd.synth(3, 0, lmin=30, lmax=60, dist_mode=2, sort_desc=0, seed=302)   # no distances, no SS
    c = set_informative_cv(cother_amd.synth(4, 0, lmin=20, lmax=90, dist_mode=0, sort_desc=0, seed=303), 304, frac=0.5)
    return cother_amd.concat([a, b, c])


def test_packed_round_trip(sets):
    pk = _packed.pack(sets)
    back = from_packed(pk)
    _same(sets, back)
    assert list(back.dists()) == list(sets.dists())


def test_query_from_the_middle_of_a_batch(sets):
    """pointers advanced past the first k profiles (PMBatchProData::PMDataSkipNPros)"""
    pk = _packed.pack(sets)
    lens = sets.lens()
    for k in (0, 3, sets.nprof - 1):
        adv = _packed.skip(pk, lens, k)
        one = from_packed(pk, nprof=1, szfields=False, advance=adv)
        _same(cother_amd.subset(sets, k, k + 1), one)


def test_malformed_buffers_refused(sets):
    L = cother_amd.lib()
    pk = _packed.pack(sets)
    # an offset table that does not match the pointers
    sz = (ctypes.c_size_t * (_packed.NFIELDS + 1))(*pk.sz)
    sz[5] += 32
    assert not L.cth_pset_from_packed(pk.beg, pk.end, sz, -1, None)
    assert b"offset" in L.cth_last_error()
    # a field shorter than the profiles need
    end = (ctypes.c_void_p * _packed.NFIELDS)(*pk.end)
    end[_packed.F_TRG + 3] -= 4
    assert not L.cth_pset_from_packed(pk.beg, end, None, -1, None)
    # a zero length
    bad = _packed.pack(sets, lens_override={2: 0})
    assert not L.cth_pset_from_packed(bad.beg, bad.end, bad.sz, -1, None)
    assert b"length" in L.cth_last_error()


@pytest.mark.skipif(not (os.path.exists(MAKEDB) and os.path.exists(DB2BIN)),
                    reason="reference database tools not built (oracle/Makefile.ref needs /root/reference)")
def test_chunk_of_a_reference_database(tmp_path, sets):
    files = []
    for p in range(sets.nprof):
        f = str(tmp_path / f"p{p}.pro")
        cother_amd.write_text_profiles(cother_amd.subset(sets, p, p + 1), f)
        files.append(f)
    db = str(tmp_path / "db")
    r = subprocess.run([MAKEDB, "-o", db, "-p", OPTIONS] + files, capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([DB2BIN, "-d", db, "-p", OPTIONS], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    whole = cother_amd.read_binary_db(db + ".tbin")
    for p0, p1 in ((0, whole.nprof), (3, 9), (whole.nprof - 2, whole.nprof)):
        pk = _packed.pack_sections(db + ".tbin", p0, p1)
        chunk = from_packed(pk)
        _same(cother_amd.subset(whole, p0, p1), chunk)


@pytest.mark.gpu
def test_packed_uploads_search_alike(ctx, sets):
    """queries and database uploaded from packed buffers search exactly as the
    same sets uploaded field by field"""
    L = cother_amd.lib()
    q = cother_amd.synth(2, 0, lmin=60, lmax=140, dist_mode=0, sort_desc=0, seed=305)
    db = cother_amd.concat([cother_amd.subset(q, 0, 2), sets,
                            cother_amd.synth(300, 0, lmin=30, lmax=250, dist_mode=1, sort_desc=1, seed=306)])
    ctx.upload_queries(q)
    ctx.upload_db(db)
    want, wn, _ = ctx.search(eval=1e3)
    pq, pd = _packed.pack(q), _packed.pack(db)
    cother_amd._check(L.cth_upload_queries_packed(ctx.h, pq.beg, pq.end, pq.sz, -1), "upload_queries_packed")
    cother_amd._check(L.cth_upload_db_packed(ctx.h, pd.beg, pd.end, pd.sz), "upload_db_packed")
    got, gn, _ = ctx.search(eval=1e3)
    print(f"{len(got)} hits (field-by-field upload {len(want)})", flush=True)
    assert len(want) >= 2 and (gn == wn).all()
    for f in want.dtype.names:
        assert (got[f] == want[f]).all(), f
