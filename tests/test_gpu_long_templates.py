"""Templates longer than the LDS holds boundary rows for (CTH_LDS_ROW_MAX =
8192 positions): phase-1 DP and the MAP kernels keep their stripe boundary
rows in global memory; results equal the oracle's bit for bit, alone and in a
database that mixes long and short profiles.  The reference accepts profiles
up to MAXCOLUMNS = 50000 positions (liblib/mylimits.h:50)."""
import ctypes

import numpy as np
import pytest

import cother_amd

pytestmark = pytest.mark.gpu

FIELDS = ("profile", "score", "psts", "beg", "end", "mapbeg", "mapend", "idnts", "ngaps", "alnlen", "pstvs_map",
          "mapscore")
STATS = ("logE", "logEpA", "lambda_est", "K_est", "minsc", "maxsc", "expected", "H", "lambda_", "K")
MAPF = ("maxsc", "end", "beg", "idnts", "psts", "ngaps", "alnlen", "logz", "maxlog")


def splice(pieces):
    """One profile made of the positions of the given one-profile sets in order
    (the first piece's background, ENO and beginning transitions)."""
    lens = np.array([sum(int(p.lens()[0]) for p in pieces)], np.int32)
    out = cother_amd.Pset(cother_amd.lib().cth_pset_alloc(1, lens.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    o = 0
    out.field("trn")[:, 0] = pieces[0].field("trn")[:, 0]
    for p in pieces:
        n = int(p.lens()[0])
        for name in ("trg", "ss", "cv"):
            out.field(name)[:, o:o + n] = p.field(name)
        for name in ("cvprior", "cvnorm2", "aa", "sss", "dd"):
            out.field(name)[o:o + n] = p.field(name)
        out.field("trn")[:, 1 + o:1 + o + n] = p.field("trn")[:, 1:]
        o += n
    out.field("bkg")[0] = pieces[0].field("bkg")[0]
    out.field("eno")[0] = pieces[0].field("eno")[0]
    return out


@pytest.fixture(scope="module")
def long_sets():
    q = cother_amd.synth(1, 2, lmin=300, lmax=300, dist_mode=0, sort_desc=0, seed=81)
    # the query inside a 9100-position template: a strong hit whose alignment
    # crosses many stripes of the long rows
    a = cother_amd.synth(1, 2, lmin=4000, lmax=4000, dist_mode=2, sort_desc=0, seed=82)
    b = cother_amd.synth(1, 2, lmin=4800, lmax=4800, dist_mode=2, sort_desc=0, seed=83)
    host = splice([a, q, b])
    other = cother_amd.synth(1, 2, lmin=8300, lmax=8300, dist_mode=2, sort_desc=0, seed=84)
    short = cother_amd.synth(40, 0, lmin=40, lmax=300, dist_mode=1, sort_desc=1, seed=85)
    db = cother_amd.concat([host, other, short])
    return q, db


def test_phase1_on_long_templates(ctx, oracle, long_sets):
    q, db = long_sets
    assert db.lens()[0] == 9100 and db.lens()[1] == 8300
    ctx.upload_queries(q)
    ctx.upload_db(db)
    res, bt = ctx.phase1_block(0, 0, 3, 0.0, want_btck=True)
    sc, _ = oracle.score_matrix(q, 0, db, 0, 3)
    ores, obt = oracle.phase1(q, 0, db, 0, 3, sc, 0.0)
    bad = np.argwhere(bt != obt)
    assert bad.size == 0, f"{len(bad)} traceback bytes differ, first {bad[:5].tolist()}"
    for f in ("maxsc", "maxxy", "passed", "psts", "beg"):
        assert (res[f] == ores[f]).all(), (f, res[f], ores[f])
    assert (res["score"].view(np.int32) == ores["score"].view(np.int32)).all()
    # the planted copy: the maximum ends in the query's segment of the template
    assert 4000 <= (int(res["maxxy"][0]) >> 16) < 4300


def test_map_on_long_templates(ctx, oracle, long_sets):
    q, db = long_sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    prof = np.array([0, 1, 2], np.int32)
    epa = np.array([-60.0, -21.0, -5.0], np.float32)
    res = ctx.map_block(0, 0, 3, prof, epa)
    sc, md = oracle.score_matrix(q, 0, db, 0, 3)
    d = db.dists()
    for j, p in enumerate(prof):
        want = oracle.map_profile(q, 0, db, int(p), sc, md, int(d[p] - d[0]), float(epa[j]))
        for f in MAPF:
            assert np.asarray(res[j][f]).tobytes() == np.asarray(want[f]).tobytes(), (p, f, res[j][f], want[f])
    assert res[0]["alnlen"] > 250


@pytest.mark.parametrize("noalloc", [False, True])
def test_search_with_long_templates(ctx, oracle, long_sets, noalloc):
    """The whole search over a database with two long profiles first (length
    order), as one chunk and, through the no-allocation entry point, in chunks
    of a 256 MiB scratch (a 300-position query against 9100 columns fits)."""
    q, db = long_sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    hits, nh, st = ctx.search(eval=1e3, noalloc=noalloc)
    want = oracle.search_query(q, 0, db, eval_=1e3)
    print(f"{len(hits)} hits (oracle {len(want)}), {st.nchunks:.0f} chunks", flush=True)
    assert len(hits) == len(want) >= 1
    assert hits[0]["profile"] == 0 and hits[0]["alnlen"] > 250
    for f in FIELDS:
        assert (hits[f] == want[f]).all(), f
    for f in STATS:
        np.testing.assert_allclose(hits[f], want[f], rtol=2e-6, atol=1e-30, err_msg=f)

