"""A search that starts from files: the query as a COTHER text profile
(cth_read_text_profiles) with its distance lists from an adddist-format
distance file (cth_add_distances), the database as a binary file written by
the reference's own makedb and db2bin (cth_read_binary_db).  The fixtures
are in tests/golden/files/ (written by tests/golden/make_files.py).

- CPU: the query profile inside the reference-written database carries the
  values the two readers produce for query.pro + query.dst (SS probabilities
  pass through the reference's floored text writer twice);
- GPU: the hits of the search equal the oracle's on the same profile sets,
  with the query's true hit first."""
import gzip
import os
import shutil

import numpy as np
import pytest

import cother_amd
from tests.test_proio import _same, _ss_written

FILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "files")
FIELDS = ("profile", "score", "psts", "beg", "end", "mapbeg", "mapend", "idnts", "ngaps", "alnlen", "pstvs_map",
          "mapscore")
STATS = ("logE", "logEpA", "lambda_est", "K_est", "minsc", "maxsc", "expected", "H", "lambda_", "K")


@pytest.fixture(scope="module")
def file_sets(tmp_path_factory):
    q = cother_amd.read_text_profiles(os.path.join(FILES, "query.pro"))
    cother_amd.add_distances(q, 0, os.path.join(FILES, "query.dst"))
    tbin = str(tmp_path_factory.mktemp("files") / "db.tbin")
    with gzip.open(os.path.join(FILES, "db.tbin.gz"), "rb") as fi, open(tbin, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    db = cother_amd.read_binary_db(tbin)
    return q, db


def _find(db, q):
    la, aa = db.lens(), db.field("aa")
    d = db.dists()
    qa = q.field("aa")[:q.lens()[0]]
    for p in range(db.nprof):
        if la[p] == q.lens()[0] and (aa[d[p]:d[p] + la[p]] == qa).all():
            return p
    return None


def test_query_in_the_reference_database(file_sets):
    q, db = file_sets
    assert db.nprof == 14
    p = _find(db, q)
    assert p is not None
    one = cother_amd.subset(db, p, p + 1)
    assert (q.field("dd")[:, 0] > 0).sum() > 50  # the distance file gave lists
    _same(q, one, skip=("ss",))
    assert (one.field("ss").view(np.int32) == _ss_written(_ss_written(q.field("ss"))).view(np.int32)).all()


@pytest.mark.gpu
def test_search_from_files_matches_oracle(ctx, oracle, file_sets):
    q, db = file_sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    hits, nh, st = ctx.search(eval=10.0)
    want = oracle.search_query(q, 0, db, eval_=10.0)
    print(f"{len(hits)} hits (oracle {len(want)}), first {hits['profile'][:3]}", flush=True)
    assert len(hits) == len(want) >= 2
    assert hits[0]["profile"] == _find(db, q) and hits[0]["alnlen"] >= 100
    for f in FIELDS:
        assert (hits[f] == want[f]).all(), f
    for f in STATS:
        np.testing.assert_allclose(hits[f], want[f], rtol=2e-6, atol=1e-30, err_msg=f)
