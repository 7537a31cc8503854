"""End to end: the results files (plain text and JSON) of a GPU search equal, byte for
byte, the file made from the oracle's hit records and the oracle's MAP
alignment rows (the writer itself is pinned by the reference's AlnWriter in
test_results_plain.py)."""
import math

import numpy as np
import pytest

import cother_amd

pytestmark = pytest.mark.gpu


def _buf(hits, rows):
    out = bytearray()
    hits = hits.copy()
    for h, r in zip(hits, rows):
        h["alnoff"] = len(out)
        for row in r:
            out += row.encode() if isinstance(row, str) else row
    return hits, bytes(out)


def test_results_file_matches_oracle(ctx, oracle, tmp_path):
    q = cother_amd.synth(3, 0, lmin=50, lmax=180, dist_mode=1, sort_desc=0, seed=91)
    same = cother_amd.synth(3, 0, lmin=50, lmax=180, dist_mode=1, sort_desc=0, seed=91)
    other = cother_amd.synth(20, 0, lmin=30, lmax=300, dist_mode=1, sort_desc=0, seed=92)
    db = cother_amd.concat([same, other])
    ctx.upload_queries(q)
    ctx.upload_db(db)
    print("results file: 3 queries vs 23 templates, seeds 91/92", flush=True)
    hits, nh, _, alns = ctx.search(eval=100.0, alignments=True)
    info = dict(dbnames=["synthdb"], dbsize=int(db.npos), ndbentries=db.nprof, logevthld=math.log(100.0))
    d = db.dists()
    off = 0
    for qi in range(q.nprof):
        gh, ga = hits[off:off + nh[qi]], alns[off:off + nh[qi]]
        off += nh[qi]
        gh, gbuf = _buf(gh, ga)
        got = cother_amd.format_results_plain(q, qi, db, gh, gbuf, qrysernr=qi, **info)
        got_json = cother_amd.format_results_plain(q, qi, db, gh, gbuf, fmt="json", qrysernr=qi, **info)
        want_hits = oracle.search_query(q, qi, db, eval_=100.0)
        sc, md = oracle.score_matrix(q, qi, db, 0, db.nprof)
        rows = []
        for h in want_hits:
            p = int(h["profile"])
            _, r = oracle.map_profile(q, qi, db, p, sc, md, int(d[p]), float(h["logEpA"]), want_aln=True)
            rows.append(r)
        wh, wbuf = _buf(want_hits, rows)
        want = cother_amd.format_results_plain(q, qi, db, wh, wbuf, qrysernr=qi, **info)
        assert got == want, f"query {qi}"
        assert got_json == cother_amd.format_results_plain(q, qi, db, wh, wbuf, fmt="json", qrysernr=qi, **info)
        assert got.count(b"\n>") == len(want_hits)
        path = cother_amd.write_results_plain(q, qi, db, gh, gbuf, str(tmp_path), qrysernr=qi, **info)
        assert open(path, "rb").read() == got
    assert len(hits) >= 3
