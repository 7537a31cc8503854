"""Serving loop (cth_server_*): requests of mixed lengths admitted by buffer
footprint and run concurrently on several slots give, query by query, the
same hits and alignment rows as cth_search alone."""
import numpy as np
import pytest

import cother_amd

pytestmark = pytest.mark.gpu


def _alone(ctx, qi):
    hits, nh, _, alns = ctx.search(qi, qi + 1, eval=10.0, alignments=True)
    return hits, alns


@pytest.fixture(scope="module")
def sets():
    # queries of mixed lengths; the first templates are copies of two queries
    q = cother_amd.synth(9, 0, lmin=20, lmax=400, dist_mode=1, sort_desc=0, seed=131)
    same = cother_amd.subset_list(q, [1, 4])
    other = cother_amd.synth(300, 1, 220.0, 0.35, 30, 1500, dist_mode=1, sort_desc=1, seed=132)
    return q, cother_amd.concat([same, other])


@pytest.mark.parametrize("nslots,budget_cells", [(1, 0), (3, 0), (4, 2_000_000)])
def test_results_alone_and_in_the_mix(ctx, sets, nslots, budget_cells):
    q, db = sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    print(f"serve: {q.nprof} queries x {db.nprof} templates, {nslots} slots", flush=True)
    want = [_alone(ctx, qi) for qi in range(q.nprof)]
    lens = q.lens()
    # budget_cells 0: room for everything; otherwise a budget that holds only
    # a few requests at once (admission must wait for footprints to free up)
    chunk = 1 << 16
    budget = int(budget_cells * 13) if budget_cells else 1 << 34
    budget = max(budget, int(lens.max()) * chunk * 13)
    srv = cother_amd.Server(ctx, nslots=nslots, mem_budget=budget, chunk_cols=chunk, alncap=1 << 22)
    try:
        order = list(range(q.nprof)) + [3, 0, 8]  # some queries twice
        tickets = [(qi, srv.submit(qi)) for qi in order]
        seen_slots = set()
        for qi, t in tickets:
            hits, alns, info = srv.wait(t)
            assert info["qidx"] == qi
            seen_slots.add(info["slot"])
            assert info["footprint"] == int(lens[qi]) * min(chunk, int(db.npos) + 64) * 13
            wh, wa = want[qi]
            assert len(hits) == len(wh), f"query {qi}"
            for f in ("profile", "score", "logE", "beg", "end", "mapbeg", "mapend", "alnlen", "idnts", "psts"):
                assert (hits[f].view(np.uint32) == wh[f].view(np.uint32)).all(), f"query {qi} field {f}"
            assert alns == wa, f"query {qi}"
        print(f"slots used: {sorted(seen_slots)}", flush=True)
    finally:
        srv.close()


@pytest.mark.parametrize("nslots,batch", [(1, 4), (2, 8)])
def test_batched_ticks_alone_and_in_the_mix(ctx, sets, nslots, batch):
    """with batch > 1 a free slot runs the requests waiting together as one
    batched search: every request's hits and rows as cth_search alone, and
    some ticks really hold several requests"""
    q, db = sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    want = [_alone(ctx, qi) for qi in range(q.nprof)]
    # footprints of the whole database (chunk_cols 0): what a batched search
    # of those queries needs
    srv = cother_amd.Server(ctx, nslots=nslots, mem_budget=1 << 34, chunk_cols=0, alncap=1 << 22, batch=batch)
    try:
        order = list(range(q.nprof)) + [3, 0, 8, 3]  # some queries twice, one three times
        tickets = [(qi, srv.submit(qi)) for qi in order]
        groups = []
        for qi, t in tickets:
            hits, alns, info = srv.wait(t)
            assert info["qidx"] == qi
            groups.append(info["group"])
            wh, wa = want[qi]
            assert len(hits) == len(wh), f"query {qi}"
            for f in ("profile", "score", "logE", "beg", "end", "mapbeg", "mapend", "alnlen", "idnts", "psts"):
                assert (hits[f].view(np.uint32) == wh[f].view(np.uint32)).all(), f"query {qi} field {f}"
            assert alns == wa, f"query {qi}"
        print(f"{len(order)} requests, tick sizes seen {sorted(set(groups))}", flush=True)
        assert max(groups) > 1 and max(groups) <= batch
    finally:
        srv.close()


def test_request_over_budget_is_refused(ctx, sets):
    q, db = sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    srv = cother_amd.Server(ctx, nslots=1, mem_budget=1 << 20, chunk_cols=1 << 16)
    try:
        with pytest.raises(RuntimeError, match="budget"):
            srv.submit(0)
    finally:
        srv.close()


def test_uploads_refused_while_serving(ctx, sets):
    q, db = sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    srv = cother_amd.Server(ctx, nslots=2, mem_budget=1 << 34, chunk_cols=1 << 16)
    try:
        with pytest.raises(RuntimeError, match="shared"):
            ctx.upload_db(db)
    finally:
        srv.close()
    ctx.upload_db(db)  # allowed again once the server is gone


@pytest.mark.parametrize("which", ["largest", "smallest"])
def test_batched_tick_failure_stays_per_request(ctx, sets, which):
    """requests whose alignment rows overflow the per-request capacity fail as
    they do alone ("alignment buffer full"), and the requests batched with
    them get their results as alone: with the capacity just under the largest
    request's rows the batch itself fits; with it at the smallest request's
    rows the batch overflows and its requests run again one by one"""
    q, db = sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    want = [_alone(ctx, qi) for qi in range(q.nprof)]
    need = [int(5 * want[qi][0]["alnlen"].astype(np.int64).sum()) for qi in range(q.nprof)]
    cap = max(need) - 1 if which == "largest" else min(n for n in need if n > 0)
    over = [n > cap for n in need]
    assert cap > 0 and any(over) and not all(over)
    if which == "smallest":
        assert sum(need) > cap * q.nprof  # the batch's shared buffer overflows too
    print(f"row bytes per query {need}, capacity {cap}", flush=True)
    srv = cother_amd.Server(ctx, nslots=1, mem_budget=1 << 34, chunk_cols=0, alncap=cap, batch=q.nprof)
    try:
        tickets = [(qi, srv.submit(qi)) for qi in range(q.nprof)]
        for qi, t in tickets:
            if over[qi]:
                with pytest.raises(RuntimeError, match="alignment buffer full"):
                    srv.wait(t)
                continue
            hits, alns, info = srv.wait(t)
            wh, wa = want[qi]
            assert len(hits) == len(wh), f"query {qi}"
            for f in ("profile", "score", "logE", "beg", "end", "mapbeg", "mapend", "alnlen", "idnts", "psts"):
                assert (hits[f].view(np.uint32) == wh[f].view(np.uint32)).all(), f"query {qi} field {f}"
            assert alns == wa, f"query {qi}"
    finally:
        srv.close()


def test_batching_window_gathers_a_burst(ctx, sets):
    """with batch_wait_us a slot that took the first request of a burst waits
    for the rest: the burst runs as one tick, every request's results as
    cth_search alone"""
    q, db = sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    want = [_alone(ctx, qi) for qi in range(q.nprof)]
    srv = cother_amd.Server(ctx, nslots=1, mem_budget=1 << 34, chunk_cols=0, alncap=1 << 22, batch=q.nprof,
                            batch_wait_us=200000)
    try:
        tickets = [(qi, srv.submit(qi)) for qi in range(q.nprof)]
        groups = []
        for qi, t in tickets:
            hits, alns, info = srv.wait(t)
            groups.append(info["group"])
            wh, wa = want[qi]
            assert len(hits) == len(wh), f"query {qi}"
            for f in ("profile", "score", "logE", "beg", "end", "mapbeg", "mapend", "alnlen", "idnts", "psts"):
                assert (hits[f].view(np.uint32) == wh[f].view(np.uint32)).all(), f"query {qi} field {f}"
            assert alns == wa, f"query {qi}"
        assert groups == [q.nprof] * q.nprof, groups
    finally:
        srv.close()


def test_batching_window_over_a_second_refused(ctx, sets):
    with pytest.raises(RuntimeError, match="batching window"):
        cother_amd.Server(ctx, nslots=1, batch=4, batch_wait_us=2000000)
