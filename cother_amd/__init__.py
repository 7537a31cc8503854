"""cother_amd: MI355X-native COTHER search hot path (Python bindings).

The product is the C-ABI library ``libcother_hip.so`` built in-tree for
gfx950 (see include/cother_hip.h); this module binds it with ctypes for the
tests and the benchmark.  Loading fails loudly when the library is missing:
there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libcother_hip.so")
ENVDIR = os.path.join(HERE, "env")

NAA, NCV, NSS, NTRN, DDSTRIDE = 20, 19, 3, 7, 128


class CthPset(ctypes.Structure):
    """Mirror of struct cth_pset (cother_amd/csrc/cth_pset.h)."""
    _fields_ = [
        ("nprof", ctypes.c_int32),
        ("npos", ctypes.c_int64),
        ("len", ctypes.POINTER(ctypes.c_int32)),
        ("dist", ctypes.POINTER(ctypes.c_int64)),
        ("eno", ctypes.POINTER(ctypes.c_float)),
        ("bkg", ctypes.POINTER(ctypes.c_float)),
        ("aa", ctypes.POINTER(ctypes.c_char)),
        ("sss", ctypes.POINTER(ctypes.c_char)),
        ("trg", ctypes.POINTER(ctypes.c_float)),
        ("trn", ctypes.POINTER(ctypes.c_float)),
        ("ss", ctypes.POINTER(ctypes.c_float)),
        ("cv", ctypes.POINTER(ctypes.c_float)),
        ("cvprior", ctypes.POINTER(ctypes.c_float)),
        ("cvnorm2", ctypes.POINTER(ctypes.c_float)),
        ("dd", ctypes.POINTER(ctypes.c_uint8)),
        ("name", ctypes.POINTER(ctypes.c_char_p)),
        ("desc", ctypes.POINTER(ctypes.c_char_p)),
    ]


class SearchOpts(ctypes.Structure):
    """Mirror of cth_search_opts (include/cother_hip.h)."""
    _fields_ = [("eval", ctypes.c_float), ("ssemodel", ctypes.c_int), ("do_map", ctypes.c_int),
                ("timing", ctypes.c_int), ("mem_budget", ctypes.c_uint64), ("minpp", ctypes.c_float),
                ("contraction", ctypes.c_int), ("alnbuf", ctypes.c_void_p), ("alncap", ctypes.c_uint64),
                ("dbsize", ctypes.c_uint64), ("ndb", ctypes.c_uint64)]


class SearchStats(ctypes.Structure):
    """Mirror of cth_search_stats (include/cother_hip.h)."""
    _fields_ = [(n, ctypes.c_double) for n in ("cells", "ms_phase1", "ms_phase2", "npassed", "nchunks",
                                               "ms_k_score", "ms_k_dp", "ms_k_stats", "ms_k_map", "map_cells")]

    def asdict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


# Mirror of cth_hit (include/cother_hip.h)
HIT = np.dtype([("query", np.int32), ("profile", np.int32), ("score", np.float32), ("logE", np.float32),
                ("logEpA", np.float32), ("lambda_est", np.float32), ("K_est", np.float32), ("minsc", np.float32),
                ("maxsc", np.float32), ("expected", np.float32), ("H", np.float32), ("lambda_", np.float32),
                ("K", np.float32), ("psts", np.int32), ("beg", np.uint32), ("end", np.uint32),
                ("mapbeg", np.uint32), ("mapend", np.uint32), ("idnts", np.int32), ("ngaps", np.int32),
                ("alnlen", np.int32), ("pstvs_map", np.int32), ("mapscore", np.float32), ("reserved", np.int32),
                ("alnoff", np.uint64)])

DPRES = np.dtype([("maxsc", np.float32), ("maxxy", np.uint32), ("passed", np.int32), ("score", np.float32),
                  ("psts", np.int32), ("beg", np.uint32)])

# Mirror of struct cth_mapres (cother_amd/csrc/cth_map.h)
MAPRES = np.dtype([("maxsc", np.float32), ("end", np.uint32), ("beg", np.uint32), ("idnts", np.int32),
                   ("psts", np.int32), ("ngaps", np.int32), ("alnlen", np.int32), ("logz", np.float32),
                   ("maxlog", np.float32)])
MINPP = 0.28

_lib = None


def lib():
    """Load libcother_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} not built; run python -m cother_amd.build or __graft_entry__.build()")
    L = ctypes.CDLL(LIBPATH)
    P = ctypes.POINTER(CthPset)
    L.cth_synth_make.restype = P
    L.cth_synth_make.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int]
    L.cth_pset_free.argtypes = [P]
    L.cth_pset_alloc.restype = P
    L.cth_pset_alloc.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]
    L.cth_last_error.restype = ctypes.c_char_p
    L.cth_last_error.argtypes = []
    L.cth_ctx_create.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]
    L.cth_ctx_destroy.argtypes = [ctypes.c_void_p]
    L.cth_upload_queries.argtypes = [ctypes.c_void_p, P]
    L.cth_upload_db.argtypes = [ctypes.c_void_p, P]
    L.cth_score_block.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_void_p, ctypes.c_void_p]
    L.cth_phase1_block.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                   ctypes.c_void_p, ctypes.c_void_p]
    L.cth_selftest_dpp.argtypes = [ctypes.c_void_p]
    L.cth_device_count.argtypes = []
    L.cth_phase2_block.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.POINTER(ctypes.c_int), ctypes.c_void_p]
    L.cth_map_block.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_void_p]
    L.cth_selftest_dpp_shl.argtypes = [ctypes.c_void_p]
    L.cth_read_text_profiles.restype = P
    L.cth_read_text_profiles.argtypes = [ctypes.c_char_p]
    L.cth_read_binary_db.restype = P
    L.cth_read_binary_db.argtypes = [ctypes.c_char_p]
    L.cth_write_text_profiles.argtypes = [P, ctypes.c_char_p, ctypes.c_int]
    L.cth_bench_copy.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    L.cth_search.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(SearchOpts),
                             ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(SearchStats)]
    L.cth_search_noalloc.argtypes = L.cth_search.argtypes
    L.cth_search_batch.argtypes = L.cth_search.argtypes
    L.cth_search_batch_list.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(SearchOpts),
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(SearchStats)]
    L.cth_batch_graph_counts.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_int64)]
    VP = ctypes.POINTER(ctypes.c_void_p)
    L.cth_pset_from_packed.restype = P
    L.cth_pset_from_packed.argtypes = [VP, VP, ctypes.POINTER(ctypes.c_size_t), ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_char_p)]
    L.cth_pset_nprof.argtypes = [P]
    L.cth_pset_npos.restype = ctypes.c_int64
    L.cth_pset_npos.argtypes = [P]
    L.cth_upload_db_packed.argtypes = [ctypes.c_void_p, VP, VP, ctypes.POINTER(ctypes.c_size_t)]
    L.cth_upload_queries_packed.argtypes = [ctypes.c_void_p, VP, VP, ctypes.POINTER(ctypes.c_size_t), ctypes.c_int]
    L.cth_pset_name.restype = ctypes.c_char_p
    L.cth_pset_name.argtypes = [P, ctypes.c_int]
    L.cth_pset_desc.restype = ctypes.c_char_p
    L.cth_pset_desc.argtypes = [P, ctypes.c_int]
    L.cth_pset_set_text.argtypes = [P, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p]
    L.cth_search_space.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                   ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
    L.cth_search_space.restype = None
    L.cth_outinfo_init.argtypes = [ctypes.POINTER(OutInfo)]
    L.cth_outinfo_init.restype = None
    L.cth_format_hit_plain.argtypes = [P, ctypes.c_int, P, ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(OutInfo),
                                       ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]
    L.cth_format_results_plain.argtypes = [P, ctypes.c_int, P, ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p,
                                           ctypes.c_uint64, ctypes.POINTER(OutInfo), ctypes.POINTER(ctypes.c_void_p),
                                           ctypes.POINTER(ctypes.c_uint64)]
    L.cth_write_results_plain.argtypes = [P, ctypes.c_int, P, ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p,
                                          ctypes.c_uint64, ctypes.POINTER(OutInfo), ctypes.c_char_p, ctypes.c_char_p,
                                          ctypes.c_int]
    L.cth_score_block_ex.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_void_p]
    L.cth_selftest_mfma.argtypes = [ctypes.POINTER(ctypes.c_int)]
    L.cth_format_hit_json.argtypes = L.cth_format_hit_plain.argtypes
    L.cth_format_results_json.argtypes = L.cth_format_results_plain.argtypes
    L.cth_write_results_json.argtypes = L.cth_write_results_plain.argtypes
    L.cth_free_text.argtypes = [ctypes.c_void_p]
    L.cth_free_text.restype = None
    L.cth_read_distances.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                     ctypes.c_float, ctypes.c_void_p]
    L.cth_add_distances.argtypes = [P, ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                    ctypes.c_float]
    L.cth_ctx_create_shared.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    L.cth_ctx_shape.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p,
                                ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64)]
    L.cth_server_opts_init.argtypes = [ctypes.POINTER(ServerOpts)]
    L.cth_server_opts_init.restype = None
    L.cth_server_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServerOpts), ctypes.POINTER(ctypes.c_void_p)]
    L.cth_server_footprint.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
    L.cth_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
    L.cth_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_uint64,
                                  ctypes.POINTER(RequestInfo)]
    L.cth_server_destroy.argtypes = [ctypes.c_void_p]
    L.cth_server_destroy.restype = None
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {lib().cth_last_error().decode()}")


class Pset:
    """A host profile set owned by the library (cth_pset)."""

    def __init__(self, ptr):
        if not ptr:
            raise MemoryError("profile set allocation failed")
        self.ptr = ptr
        self.s = ptr.contents

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().cth_pset_free(self.ptr)
            self.ptr = None

    @property
    def nprof(self):
        return self.s.nprof

    @property
    def npos(self):
        return self.s.npos

    def lens(self):
        return np.ctypeslib.as_array(self.s.len, (self.nprof,)).copy()

    def dists(self):
        return np.ctypeslib.as_array(self.s.dist, (self.nprof,)).copy()

    def field(self, name):
        n, p = self.npos, self.nprof
        shapes = {
            "trg": (NAA, n), "ss": (NSS, n), "cv": (NCV, n), "cvprior": (n,), "cvnorm2": (n,),
            "trn": (NTRN, n + p), "bkg": (p, NAA), "eno": (p,), "dd": (n, DDSTRIDE),
        }
        if name in ("aa", "sss"):
            return np.ctypeslib.as_array(ctypes.cast(getattr(self.s, name), ctypes.POINTER(ctypes.c_uint8)), (n,))
        return np.ctypeslib.as_array(getattr(self.s, name), shapes[name])


def synth(nprof, len_mode, median=0.0, sigma=0.0, lmin=0, lmax=0, dist_mode=1, sort_desc=1, seed=1, threads=None):
    """Seeded synthetic profile set (see cother_amd/csrc/synth.cpp)."""
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    return Pset(lib().cth_synth_make(nprof, len_mode, median, sigma, lmin, lmax, dist_mode, sort_desc, seed,
                                     threads))


def read_text_profiles(path):
    """Profiles of a text profile file or makedb text database (reference format)."""
    p = lib().cth_read_text_profiles(path.encode())
    if not p:
        raise RuntimeError(lib().cth_last_error().decode())
    return Pset(p)


def read_binary_db(path):
    """A binary profile database written by the reference's db2bin."""
    p = lib().cth_read_binary_db(path.encode())
    if not p:
        raise RuntimeError(lib().cth_last_error().decode())
    return Pset(p)


def write_text_profiles(ps, path, scale=10000):
    """Write a profile set in the reference's text profile format."""
    _check(lib().cth_write_text_profiles(ps.ptr, path.encode(), scale), "cth_write_text_profiles")


def subset_list(ps, idx):
    """A profile set of the profiles idx of ps, in that order."""
    idx = [int(i) for i in idx]
    lens = ps.lens()[idx].astype(np.int32)
    out = Pset(lib().cth_pset_alloc(len(lens), lens.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    d, ln = ps.dists(), ps.lens()
    od = out.dists()
    for k, p in enumerate(idx):
        a, b = int(d[p]), int(d[p] + ln[p])
        o = int(od[k])
        for name in ("trg", "ss", "cv"):
            out.field(name)[:, o:o + (b - a)] = ps.field(name)[:, a:b]
        for name in ("cvprior", "cvnorm2", "aa", "sss"):
            out.field(name)[o:o + (b - a)] = ps.field(name)[a:b]
        out.field("dd")[o:o + (b - a)] = ps.field("dd")[a:b]
        out.field("trn")[:, o + k:o + k + (b - a) + 1] = ps.field("trn")[:, a + p:b + p + 1]
        out.field("bkg")[k] = ps.field("bkg")[p]
        out.field("eno")[k] = ps.field("eno")[p]
        set_text(out, k, pset_name(ps, p), pset_desc(ps, p))
    return out


def subset(ps, p0, p1):
    """A profile set of the profiles [p0,p1) of ps."""
    return subset_list(ps, range(p0, p1))


def concat(sets):
    """One profile set holding the profiles of `sets` in order."""
    lens = np.concatenate([s_.lens() for s_ in sets]).astype(np.int32)
    out = Pset(lib().cth_pset_alloc(len(lens), lens.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    for name in ("trg", "ss", "cv", "trn"):
        out.field(name)[...] = np.concatenate([s_.field(name) for s_ in sets], axis=1)
    for name in ("cvprior", "cvnorm2", "eno", "aa", "sss"):
        out.field(name)[...] = np.concatenate([s_.field(name) for s_ in sets])
    for name in ("bkg", "dd"):
        out.field(name)[...] = np.concatenate([s_.field(name) for s_ in sets], axis=0)
    k = 0
    for s_ in sets:
        for p in range(s_.nprof):
            set_text(out, k, pset_name(s_, p), pset_desc(s_, p))
            k += 1
    return out


class OutInfo(ctypes.Structure):
    """Mirror of cth_outinfo (include/cother_hip.h)."""
    _fields_ = [("qname", ctypes.c_char_p), ("qdesc", ctypes.c_char_p), ("qrysernr", ctypes.c_int32),
                ("logevthld", ctypes.c_float), ("dbnames", ctypes.POINTER(ctypes.c_char_p)),
                ("ndbnames", ctypes.c_int32), ("dbsize", ctypes.c_uint64), ("ndbentries", ctypes.c_uint64),
                ("nohits", ctypes.c_int32), ("noalns", ctypes.c_int32), ("dsclen", ctypes.c_int32),
                ("dscwidth", ctypes.c_int32), ("alnwidth", ctypes.c_int32), ("show", ctypes.c_int32),
                ("ssswgt", ctypes.c_float), ("ref_lambda", ctypes.c_float), ("ref_k", ctypes.c_float),
                ("gap_lambda", ctypes.c_float), ("gap_k", ctypes.c_float), ("progname", ctypes.c_char_p),
                ("progversion", ctypes.c_char_p)]


def outinfo(**kw):
    """A cth_outinfo with the reference's defaults (cth_outinfo_init) and the given fields;
    dbnames is a list of str, qname/qdesc str."""
    o = OutInfo()
    lib().cth_outinfo_init(ctypes.byref(o))
    keep = []
    for k, v in kw.items():
        if k == "dbnames":
            arr = (ctypes.c_char_p * max(1, len(v)))(*[s.encode() for s in v])
            keep.append(arr)
            o.dbnames = ctypes.cast(arr, ctypes.POINTER(ctypes.c_char_p))
            o.ndbnames = len(v)
        elif k in ("qname", "qdesc", "progname", "progversion"):
            b = v.encode() if v is not None else None
            keep.append(b)
            setattr(o, k, b)
        else:
            setattr(o, k, v)
    o._keep = keep
    return o


def set_text(ps, k, name, desc):
    """Set the name and description of profile k."""
    _check(lib().cth_pset_set_text(ps.ptr, k, None if name is None else name.encode(),
                                   None if desc is None else desc.encode()), "cth_pset_set_text")


def pset_name(ps, k):
    v = lib().cth_pset_name(ps.ptr, k)
    return None if v is None else v.decode()


def pset_desc(ps, k):
    v = lib().cth_pset_desc(ps.ptr, k)
    return None if v is None else v.decode()


def read_distances(path, prolen, columns=(3,), prbthrhld=0.0):
    """Per-position distance lists (float32 arrays) of a distance file, as adddist compiles them."""
    cols = (ctypes.c_int * max(1, len(columns)))(*columns)
    out = np.zeros((max(1, prolen), 100), np.float32)
    _check(lib().cth_read_distances(path.encode(), prolen, cols, len(columns), prbthrhld, out.ctypes.data),
           "cth_read_distances")
    return [out[m, 1:1 + int(out[m, 0])].copy() for m in range(prolen)]


def add_distances(ps, k, path, columns=(3,), prbthrhld=0.0):
    """Replace profile k's distance lists by those of a distance file."""
    cols = (ctypes.c_int * max(1, len(columns)))(*columns)
    _check(lib().cth_add_distances(ps.ptr, k, path.encode(), cols, len(columns), prbthrhld), "cth_add_distances")

def selftest_mfma():
    """Number of mismatching accumulators of the MFMA lane-layout check (0 expected)."""
    bad = ctypes.c_int()
    _check(lib().cth_selftest_mfma(ctypes.byref(bad)), "cth_selftest_mfma")
    return bad.value


def search_space(qlen, dbsize, ndb):
    """(effective search space, length adjustment) of a query of qlen positions."""
    ss, dl = ctypes.c_uint64(), ctypes.c_uint64()
    lib().cth_search_space(qlen, dbsize, ndb, ctypes.byref(ss), ctypes.byref(dl))
    return ss.value, dl.value


def _take_text(p):
    try:
        return ctypes.string_at(p)
    finally:
        lib().cth_free_text(p)


def format_hit_plain(qs, qidx, db, hit, rows, fmt="plain", **info):
    """(annotation, alignment record) bytes of one hit; rows: the five alignment rows;
    fmt "plain" or "json"."""
    o = outinfo(**info)
    h = np.array(hit, HIT).reshape(1)
    buf = b"".join(r if isinstance(r, bytes) else r.encode() for r in rows)
    a, r = ctypes.c_void_p(), ctypes.c_void_p()
    fn = lib().cth_format_hit_json if fmt == "json" else lib().cth_format_hit_plain
    _check(fn(qs.ptr, qidx, db.ptr, h.ctypes.data, buf, ctypes.byref(o), ctypes.byref(a), ctypes.byref(r)),
           "cth_format_hit_" + fmt)
    return _take_text(a), _take_text(r)


def format_results_plain(qs, qidx, db, hits, alnbuf, fmt="plain", **info):
    """The reference's results file for query qidx (fmt "plain" or "json"), as bytes."""
    o = outinfo(**info)
    h = np.ascontiguousarray(hits, HIT)
    t, n = ctypes.c_void_p(), ctypes.c_uint64()
    fn = lib().cth_format_results_json if fmt == "json" else lib().cth_format_results_plain
    _check(fn(qs.ptr, qidx, db.ptr, h.ctypes.data, len(h), alnbuf, len(alnbuf), ctypes.byref(o), ctypes.byref(t),
              ctypes.byref(n)), "cth_format_results_" + fmt)
    try:
        return ctypes.string_at(t, n.value)
    finally:
        lib().cth_free_text(t)


def write_results_plain(qs, qidx, db, hits, alnbuf, outdir, fmt="plain", **info):
    """Write the results file of query qidx into outdir (fmt "plain" or "json"); returns its path."""
    o = outinfo(**info)
    h = np.ascontiguousarray(hits, HIT)
    path = ctypes.create_string_buffer(4096)
    fn = lib().cth_write_results_json if fmt == "json" else lib().cth_write_results_plain
    _check(fn(qs.ptr, qidx, db.ptr, h.ctypes.data, len(h), alnbuf, len(alnbuf), ctypes.byref(o), outdir.encode(),
              path, len(path)), "cth_write_results_" + fmt)
    return path.value.decode()

class Context:
    """A device context (cth_ctx): tables, uploaded queries, resident DB."""

    def __init__(self, device=0, envdir=ENVDIR):
        L = lib()
        h = ctypes.c_void_p()
        _check(L.cth_ctx_create(device, envdir.encode(), ctypes.byref(h)), "cth_ctx_create")
        self.h = h

    def graph_counts(self):
        """(captures, replays) of the batched search's graph on this context."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().cth_batch_graph_counts(self.h, ctypes.byref(a), ctypes.byref(b)), "cth_batch_graph_counts")
        return a.value, b.value

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().cth_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def upload_queries(self, qs):
        _check(lib().cth_upload_queries(self.h, qs.ptr), "cth_upload_queries")
        self.qs = qs

    def upload_db(self, db):
        _check(lib().cth_upload_db(self.h, db.ptr), "cth_upload_db")
        self.db = db

    def score_block(self, qidx, p0, p1, contraction=0):
        """Score and modular-score matrices [Lq][ncols]; contraction 1/2: log-odds term from the
        BF16/FP16 matrix-core product (0: FP32, the reference's)."""
        lq = int(self.qs.lens()[qidx])
        d, ln = self.db.dists(), self.db.lens()
        ncols = int(d[p1 - 1] + ln[p1 - 1] - d[p0])
        sc = np.empty((lq, ncols), np.float32)
        md = np.empty((lq, ncols), np.float32)
        _check(lib().cth_score_block_ex(self.h, qidx, p0, p1, contraction, sc.ctypes.data, md.ctypes.data),
               "cth_score_block_ex")
        return sc, md

    def phase1_block(self, qidx, p0, p1, scorethld, want_btck=False):
        """Phase-1 records (DPRES) of query qidx vs profiles [p0,p1), optionally the traceback bytes."""
        res = np.zeros(p1 - p0, DPRES)
        bt = None
        if want_btck:
            lq = int(self.qs.lens()[qidx])
            d, ln = self.db.dists(), self.db.lens()
            ncols = int(d[p1 - 1] + ln[p1 - 1] - d[p0])
            bt = np.zeros((lq, ncols), np.uint8)
        _check(lib().cth_phase1_block(self.h, qidx, p0, p1, scorethld, res.ctypes.data,
                                      bt.ctypes.data if bt is not None else None), "cth_phase1_block")
        return res, bt

    def phase2_block(self, qidx, p0, p1, scorethld, searchspace, ssemodel=2):
        """Phase 1 plus statistics: (records, passed indices relative to p0, stats [npass,10])."""
        n = p1 - p0
        res = np.zeros(n, DPRES)
        plist = np.zeros(n, np.int32)
        stats = np.zeros((n, 10), np.float64)
        npass = ctypes.c_int()
        _check(lib().cth_phase2_block(self.h, qidx, p0, p1, scorethld, searchspace, ssemodel, res.ctypes.data,
                                      plist.ctypes.data, ctypes.byref(npass), stats.ctypes.data),
               "cth_phase2_block")
        k = npass.value
        return res, plist[:k], stats[:k]

    def map_block(self, qidx, p0, p1, profiles, logepa, minpp=MINPP):
        """MAP stage records (MAPRES) of query qidx for the given profiles of block [p0,p1)."""
        prof = np.ascontiguousarray(profiles, np.int32)
        epa = np.ascontiguousarray(logepa, np.float32)
        res = np.zeros(len(prof), MAPRES)
        _check(lib().cth_map_block(self.h, qidx, p0, p1, prof.ctypes.data, epa.ctypes.data, len(prof), minpp,
                                   res.ctypes.data), "cth_map_block")
        return res

    def search(self, q0=0, q1=None, eval=10.0, ssemodel=2, do_map=True, timing=False, mem_budget=0,
               maxhits=None, minpp=MINPP, alignments=False, contraction=0, dbsize=0, ndb=0, noalloc=False,
               batch=False):
        """Search queries [q0,q1) against the resident database.

        dbsize, ndb: positions and profiles of the whole database when the resident set is one
        part of it (0, 0: the resident set's own).  noalloc: run cth_search_noalloc, whose
        buffers come from the library's fixed per-device scratch array (mem_budget unused).
        batch: run cth_search_batch (many queries per launch, graph-replayed).

        Returns (hits HIT array grouped by query, per-query hit counts, SearchStats), and with
        alignments=True also a list of five-row alignments (query, middle, target, query SS,
        target SS) per hit."""
        if q1 is None:
            q1 = self.qs.nprof
        if maxhits is None:
            maxhits = max(1, (q1 - q0) * self.db.nprof)
        hits = np.zeros(maxhits, HIT)
        nh = np.zeros(q1 - q0, np.int32)
        st = SearchStats()
        abuf = None
        if alignments:
            cap = 5 * min(maxhits, 4096) * int(self.qs.lens()[q0:q1].max() + self.db.lens().max())
            abuf = ctypes.create_string_buffer(cap)
        opts = SearchOpts(eval, ssemodel, int(do_map), int(timing), mem_budget, minpp, contraction,
                          ctypes.cast(abuf, ctypes.c_void_p) if abuf is not None else None,
                          len(abuf) if abuf is not None else 0, dbsize, ndb)
        name = "cth_search_noalloc" if noalloc else ("cth_search_batch" if batch else "cth_search")
        fn = getattr(lib(), name)
        _check(fn(self.h, q0, q1, ctypes.byref(opts), hits.ctypes.data, maxhits, nh.ctypes.data, ctypes.byref(st)),
               name)
        hits = hits[:int(nh.sum())]
        if not alignments:
            return hits, nh, st
        raw = abuf.raw
        alns = []
        for h in hits:
            n, o = int(h["alnlen"]), int(h["alnoff"])
            alns.append(tuple(raw[o + r * n:o + (r + 1) * n].decode() for r in range(5)))
        return hits, nh, st, alns

    def search_list(self, queries, eval=10.0, ssemodel=2, do_map=True, timing=False, mem_budget=0, minpp=MINPP,
                    alignments=False):
        """cth_search_batch_list: the batched search for the distinct query indices `queries`, in that
        order; returns (hits grouped by query in list order, per-query hit counts, SearchStats) and with
        alignments=True the five-row alignments per hit."""
        ql = np.ascontiguousarray(queries, np.int32)
        maxhits = max(1, len(ql) * self.db.nprof)
        hits = np.zeros(maxhits, HIT)
        nh = np.zeros(len(ql), np.int32)
        st = SearchStats()
        abuf = None
        if alignments:
            cap = 5 * min(maxhits, 4096) * int(self.qs.lens()[ql].max() + self.db.lens().max())
            abuf = ctypes.create_string_buffer(cap)
        opts = SearchOpts(eval, ssemodel, int(do_map), int(timing), mem_budget, minpp, 0,
                          ctypes.cast(abuf, ctypes.c_void_p) if abuf is not None else None,
                          len(abuf) if abuf is not None else 0, 0, 0)
        _check(lib().cth_search_batch_list(self.h, ql.ctypes.data, len(ql), ctypes.byref(opts), hits.ctypes.data,
                                           maxhits, nh.ctypes.data, ctypes.byref(st)), "cth_search_batch_list")
        hits = hits[:int(nh.sum())]
        if not alignments:
            return hits, nh, st
        raw = abuf.raw
        alns = [tuple(raw[int(h["alnoff"]) + r * int(h["alnlen"]):int(h["alnoff"]) + (r + 1) * int(h["alnlen"])]
                      .decode() for r in range(5)) for h in hits]
        return hits, nh, st, alns


class ServerOpts(ctypes.Structure):
    """Mirror of cth_server_opts (include/cother_hip.h)."""
    _fields_ = [("nslots", ctypes.c_int), ("mem_budget", ctypes.c_uint64), ("chunk_cols", ctypes.c_uint64),
                ("search", SearchOpts), ("alncap", ctypes.c_uint64), ("batch", ctypes.c_int),
                ("batch_wait_us", ctypes.c_uint32)]


class RequestInfo(ctypes.Structure):
    """Mirror of cth_request_info (include/cother_hip.h)."""
    _fields_ = [("qidx", ctypes.c_int32), ("slot", ctypes.c_int32), ("footprint", ctypes.c_uint64),
                ("ms_queued", ctypes.c_double), ("ms_latency", ctypes.c_double), ("ms_service", ctypes.c_double),
                ("stats", SearchStats), ("group", ctypes.c_int32)]


class Server:
    """The serving loop (cth_server_*) over a context's uploaded queries and resident database."""

    def __init__(self, ctx, nslots=2, mem_budget=64 << 30, chunk_cols=4 << 20, eval=10.0, ssemodel=2, do_map=True,
                 minpp=MINPP, alncap=0, contraction=0, batch=1, batch_wait_us=0):
        """batch: requests a slot may take in one tick (run together by cth_search_batch_list);
        batch_wait_us: how long a slot waits from its first request for more (0: no wait)."""
        o = ServerOpts()
        lib().cth_server_opts_init(ctypes.byref(o))
        o.nslots, o.mem_budget, o.chunk_cols, o.alncap = nslots, mem_budget, chunk_cols, alncap
        o.batch, o.batch_wait_us = batch, batch_wait_us
        o.search.eval, o.search.ssemodel, o.search.do_map, o.search.minpp = eval, ssemodel, int(do_map), minpp
        o.search.contraction = contraction
        self.ctx = ctx
        self.ndb = ctx.db.nprof
        self.alncap = alncap
        h = ctypes.c_void_p()
        _check(lib().cth_server_create(ctx.h, ctypes.byref(o), ctypes.byref(h)), "cth_server_create")
        self.h = h

    def footprint(self, qidx):
        b = ctypes.c_uint64()
        _check(lib().cth_server_footprint(self.h, qidx, ctypes.byref(b)), "cth_server_footprint")
        return b.value

    def submit(self, qidx):
        t = ctypes.c_uint64()
        _check(lib().cth_server_submit(self.h, qidx, ctypes.byref(t)), "cth_server_submit")
        return t.value

    def wait(self, ticket):
        """(hits, alignments or None, info dict) of a submitted request."""
        hits = np.zeros(max(1, self.ndb), HIT)
        n = ctypes.c_int()
        buf = ctypes.create_string_buffer(self.alncap) if self.alncap else None
        info = RequestInfo()
        _check(lib().cth_server_wait(self.h, ticket, hits.ctypes.data, len(hits), ctypes.byref(n), buf,
                                     self.alncap, ctypes.byref(info)), "cth_server_wait")
        hits = hits[:n.value]
        alns = None
        if buf is not None:
            raw = buf.raw
            alns = [tuple(raw[int(h["alnoff"]) + r * int(h["alnlen"]):int(h["alnoff"]) + (r + 1) * int(h["alnlen"])]
                          .decode() for r in range(5)) for h in hits]
        d = {k: getattr(info, k) for k, _ in RequestInfo._fields_ if k != "stats"}
        d["stats"] = info.stats.asdict()
        return hits, alns, d

    def close(self):
        if getattr(self, "h", None):
            lib().cth_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
