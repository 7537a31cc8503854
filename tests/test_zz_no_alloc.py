"""cth_search_noalloc (buffers from the library's fixed per-device scratch
array, no device allocation) against cth_search on the same inputs: hits,
statistics and alignment rows identical, over two calls in a row (the second
with a smaller input, so contents the first call left in the scratch cannot
leak into it), no device memory taken by the calls, and a query too large for
the scratch refused with an error."""
import ctypes

import numpy as np
import pytest

import cother_amd

pytestmark = pytest.mark.gpu

SCRATCH = 256 << 20  # CTH_NOALLOC_SCRATCH_BYTES


def free_device_bytes():
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def same_hits(a, b, alns_a, alns_b):
    assert len(a) == len(b)
    for f in a.dtype.names:
        if f == "alnoff":
            continue
        assert (a[f].view(np.uint32) == b[f].view(np.uint32)).all() if a[f].dtype.itemsize == 4 \
            else (a[f] == b[f]).all(), f
    assert alns_a == alns_b


def test_noalloc_equals_search_on_two_calls():
    if cother_amd.lib().cth_device_count() < 1:
        pytest.skip("no GPU")
    # three queries up to 420 positions against 1500 templates (~2.7e5
    # positions): the scratch's chunk share (192 MiB) holds ~3.9e4 columns of
    # a 420-position query, so the long queries run in several chunks
    q = cother_amd.synth(3, 0, lmin=250, lmax=420, dist_mode=0, sort_desc=0, seed=71)
    rest = cother_amd.synth(1500, 0, lmin=60, lmax=300, dist_mode=1, sort_desc=1, seed=72)
    db = cother_amd.concat([cother_amd.subset(q, 0, 2), rest])  # two self-hits with long alignments
    na = cother_amd.Context(0)
    ref = cother_amd.Context(0)
    try:
        for c in (na, ref):
            c.upload_queries(q)
            c.upload_db(db)
        print(f"noalloc: {q.nprof} queries (lengths {list(q.lens())}) vs {db.nprof} templates "
              f"({db.npos} positions)", flush=True)
        h1, n1, s1, a1 = na.search(0, 3, eval=1e3, alignments=True, noalloc=True)
        r1, m1, t1, b1 = ref.search(0, 3, eval=1e3, alignments=True)
        print(f"call 1: {len(h1)} hits (search {len(r1)}), {s1.nchunks:.0f} chunks (search {t1.nchunks:.0f})",
              flush=True)
        assert s1.nchunks > t1.nchunks >= q.nprof  # the scratch splits the database more finely
        assert len(r1) >= 3 and (n1 == m1).all()
        same_hits(h1, r1, a1, b1)
        # second call: a smaller input (the shortest query alone) on the same scratch
        short = int(np.argmin(q.lens()))
        free_before = free_device_bytes()
        h2, n2, s2, a2 = na.search(short, short + 1, eval=1e3, alignments=True, noalloc=True)
        assert free_device_bytes() == free_before  # nothing allocated on the device
        r2, m2, t2, b2 = ref.search(short, short + 1, eval=1e3, alignments=True)
        print(f"call 2 (query {short}): {len(h2)} hits (search {len(r2)})", flush=True)
        assert len(r2) >= 1 and (n2 == m2).all()
        same_hits(h2, r2, a2, b2)
        # the same two calls without MAP (the other chunk layout)
        for (lo, hi) in ((0, 3), (short, short + 1)):
            h, n, s = na.search(lo, hi, eval=1e3, do_map=False, noalloc=True)
            r, m, t = ref.search(lo, hi, eval=1e3, do_map=False)
            same_hits(h, r, None, None)
    finally:
        na.close()
        ref.close()


def test_noalloc_refuses_a_query_over_the_scratch():
    if cother_amd.lib().cth_device_count() < 1:
        pytest.skip("no GPU")
    # 4000 query positions x one 5000-position template x 13 bytes = 260 MB,
    # over the 192 MiB the chunk buffers may take: refused before any kernel
    q = cother_amd.synth(1, 2, lmin=4000, lmax=4000, dist_mode=2, sort_desc=0, seed=73)
    db = cother_amd.synth(1, 2, lmin=5000, lmax=5000, dist_mode=2, sort_desc=0, seed=74)
    small = cother_amd.synth(1, 2, lmin=50, lmax=50, dist_mode=2, sort_desc=0, seed=75)
    c = cother_amd.Context(0)
    try:
        c.upload_queries(cother_amd.concat([q, small]))
        c.upload_db(db)
        assert 4000 * 5056 * 13 > SCRATCH // 4 * 3
        with pytest.raises(RuntimeError, match="scratch allows"):
            c.search(0, 1, noalloc=True)
        # a query that fits still runs on the same context after the refusal
        h, n, s = c.search(1, 2, noalloc=True)
        r, m, t = c.search(1, 2)
        same_hits(h, r, None, None)
    finally:
        c.close()
