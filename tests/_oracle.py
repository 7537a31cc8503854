"""ctypes bindings of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: it computes the expected values the HIP
path is checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ODIR = os.path.join(ROOT, "oracle")
OLIB = os.path.join(ODIR, "liboracle.so")
REFLIB = os.path.join(ODIR, "_ref", "libddms_ref.so")


class OrcTables(ctypes.Structure):
    _fields_ = [
        ("ss", ctypes.c_float * 467),
        ("cvcard", ctypes.c_int),
        ("cvshift", ctypes.c_int),
        ("cvkeys", ctypes.c_float * 128),
        ("cvvals", ctypes.c_float * 128),
        ("cvkappa0inv", ctypes.c_float),
        ("cvpowernu0", ctypes.c_float),
        ("cvcterm", ctypes.c_float),
        ("dms", ctypes.c_float * (22 * 22)),
    ]


class Oracle:
    def __init__(self, L):
        self.L = L
        vp = ctypes.c_void_p
        L.orc_logf.restype = ctypes.c_float
        L.orc_logf.argtypes = [ctypes.c_float]
        L.orc_tables_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(OrcTables)]
        L.orc_cell_init.restype = ctypes.c_float
        L.orc_cell_init.argtypes = [vp, ctypes.c_int64, ctypes.c_int, vp, ctypes.c_int64, ctypes.c_int]
        for f in ("orc_cell_ss", "orc_cell_cvs"):
            getattr(L, f).restype = ctypes.c_float
            getattr(L, f).argtypes = [ctypes.POINTER(OrcTables), vp, ctypes.c_int64, vp, ctypes.c_int64]
        L.orc_cell_ddms.restype = ctypes.c_float
        L.orc_cell_ddms.argtypes = [ctypes.POINTER(OrcTables), vp, vp]
        L.orc_ddms_tab.restype = ctypes.c_float
        L.orc_ddms_tab.argtypes = [vp, vp, vp, vp, vp]
        L.orc_score_matrix.argtypes = [ctypes.POINTER(OrcTables), vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int,
                                       vp, vp, ctypes.c_int]
        L.orc_dp_phase1.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_float,
                                    vp, vp, vp, vp, vp, vp, ctypes.c_int]
        L.orc_score_threshold.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_uint64,
                                          ctypes.c_uint64, ctypes.POINTER(ctypes.c_float),
                                          ctypes.POINTER(ctypes.c_float)]
        L.orc_stat_params.argtypes = [vp, vp]
        L.orc_lbetainc.restype = ctypes.c_double
        L.orc_lbetainc.argtypes = [ctypes.c_double] * 3
        L.orc_lgammaincq.restype = ctypes.c_double
        L.orc_lgammaincq.argtypes = [ctypes.c_double] * 2
        L.orc_significance.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 7 + [vp]
        L.orc_histogram.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int, vp, vp, vp]
        L.orc_length_adjustment.restype = ctypes.c_uint64
        L.orc_length_adjustment.argtypes = [ctypes.c_float] * 4 + [ctypes.c_uint64] * 3 + [ctypes.POINTER(ctypes.c_uint64)]
        L.orc_search_query.restype = ctypes.c_int
        L.orc_search_query.argtypes = [ctypes.POINTER(OrcTables), vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
                                       vp, ctypes.c_int, ctypes.c_int]
        L.orc_map_profile.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, vp, vp, ctypes.c_int64, ctypes.c_int64,
                                      ctypes.c_float, ctypes.c_float, vp, vp, ctypes.c_int]
        self.tables = OrcTables()
        from cother_amd import ENVDIR
        rc = L.orc_tables_load(ENVDIR.encode(), ctypes.byref(self.tables))
        if rc:
            raise RuntimeError(f"oracle tables failed: {rc}")

    def logf(self, x):
        return self.L.orc_logf(x)

    def score_matrix(self, q, qp, db, p0, p1, threads=None):
        lq = int(q.lens()[qp])
        d, ln = db.dists(), db.lens()
        ncols = int(d[p1 - 1] + ln[p1 - 1] - d[p0])
        sc = np.empty((lq, ncols), np.float32)
        md = np.empty((lq, ncols), np.float32)
        if threads is None:
            threads = min(16, os.cpu_count() or 1)
        self.L.orc_score_matrix(ctypes.byref(self.tables), ctypes.cast(q.ptr, ctypes.c_void_p), qp,
                                ctypes.cast(db.ptr, ctypes.c_void_p), p0, p1, sc.ctypes.data, md.ctypes.data,
                                threads)
        return sc, md

    def phase1(self, q, qp, db, p0, p1, scores, scorethld, threads=None):
        """DP + finalisation; returns (records as DPRES array, btck)."""
        from cother_amd import DPRES
        n = p1 - p0
        lq, ncols = scores.shape
        bt = np.zeros((lq, ncols), np.uint8)
        maxsc = np.zeros(n, np.float32); maxxy = np.zeros(n, np.uint32); passed = np.zeros(n, np.int32)
        score = np.zeros(n, np.float32); psts = np.zeros(n, np.int32); beg = np.zeros(n, np.uint32)
        if threads is None:
            threads = min(16, os.cpu_count() or 1)
        sc = np.ascontiguousarray(scores, np.float32)
        self.L.orc_dp_phase1(ctypes.cast(q.ptr, ctypes.c_void_p), qp, ctypes.cast(db.ptr, ctypes.c_void_p), p0, p1,
                             sc.ctypes.data, bt.ctypes.data, scorethld, maxsc.ctypes.data, maxxy.ctypes.data,
                             passed.ctypes.data, score.ctypes.data, psts.ctypes.data, beg.ctypes.data, threads)
        res = np.zeros(n, DPRES)
        res["maxsc"], res["maxxy"], res["passed"] = maxsc, maxxy, passed
        res["score"], res["psts"], res["beg"] = score, psts, beg
        return res, bt

    def score_threshold(self, eval_, nqyposs, dbsize, lam=0.561132, K=0.001488):
        s = ctypes.c_float(); le = ctypes.c_float()
        self.L.orc_score_threshold(eval_, lam, K, nqyposs, dbsize, ctypes.byref(s), ctypes.byref(le))
        return s.value, le.value

    def stat_params(self, counts):
        c = np.ascontiguousarray(counts, np.float64)
        out = np.zeros(6, np.float64)
        self.L.orc_stat_params(c.ctypes.data, out.ctypes.data)
        return out

    def significance(self, ssemodel, lq, lt, qeno, teno, sspace, score, psts, lam, K):
        out = np.zeros(4, np.float64)
        self.L.orc_significance(ssemodel, lq, lt, qeno, teno, sspace, score, psts, lam, K, out.ctypes.data)
        return out

    def length_adjustment(self, m, n, N, lam=0.561132, K=0.001488, alpha=0.5, beta=-1.0):
        dl = ctypes.c_uint64()
        ss = self.L.orc_length_adjustment(lam, K, alpha, beta, m, n, N, ctypes.byref(dl))
        return ss, dl.value

    def search_query(self, q, qp, db, p0=0, p1=None, eval_=10.0, ssemodel=2, dbsize=None, ndb=None,
                     threads=None, do_map=True):
        """Hits of query qp against profiles [p0,p1) (HIT records sorted by logE, profile)."""
        from cother_amd import HIT
        if p1 is None:
            p1 = db.nprof
        if dbsize is None:
            dbsize = db.npos
        if ndb is None:
            ndb = db.nprof
        if threads is None:
            threads = min(16, os.cpu_count() or 1)
        hits = np.zeros(max(1, p1 - p0), HIT)
        n = self.L.orc_search_query(ctypes.byref(self.tables), ctypes.cast(q.ptr, ctypes.c_void_p), qp,
                                    ctypes.cast(db.ptr, ctypes.c_void_p), p0, p1, eval_, ssemodel, int(do_map), dbsize, ndb,
                                    hits.ctypes.data, len(hits), threads)
        if n < 0:
            raise RuntimeError("orc_search_query: hit buffer too small")
        return hits[:n]

    def map_profile(self, q, qp, db, p, sc, md, col, logepa, minpp=0.28, want_aln=False):
        """MAP stage of query qp vs profile p on the score/modular-score block
        sc, md ([Lq][ncols], profile p at column col): a MAPRES record (and the
        five alignment rows: query, middle, target, query SS, target SS)."""
        from cother_amd import MAPRES
        res = np.zeros(1, MAPRES)
        cap = int(q.lens()[qp]) + int(db.lens()[p]) + 2
        aln = ctypes.create_string_buffer(5 * cap) if want_aln else None
        s_ = np.ascontiguousarray(sc, np.float32)
        m_ = np.ascontiguousarray(md, np.float32)
        self.L.orc_map_profile(ctypes.cast(q.ptr, ctypes.c_void_p), qp, ctypes.cast(db.ptr, ctypes.c_void_p), p,
                               s_.ctypes.data, m_.ctypes.data, s_.shape[1], col, logepa, minpp, res.ctypes.data,
                               aln, cap)
        if not want_aln:
            return res[0]
        n = int(res[0]["alnlen"])
        raw = aln.raw
        return res[0], tuple(raw[r * cap:r * cap + n].decode() for r in range(5))

    def ddms_tab(self, tab22, qrec, trec):
        tab = np.ascontiguousarray(tab22, np.float32)
        q = np.ascontiguousarray(qrec, np.uint8)
        t = np.ascontiguousarray(trec, np.uint8)
        m = np.zeros(3, np.float32)
        c = np.zeros(3, np.uint32)
        v = self.L.orc_ddms_tab(tab.ctypes.data, q.ctypes.data, t.ctypes.data, m.ctypes.data, c.ctypes.data)
        return v, m, c


_cache = None


def load():
    global _cache
    if _cache is None:
        if not os.path.exists(OLIB):
            subprocess.run(["make", "-s", "-C", ODIR], check=True)
        _cache = Oracle(ctypes.CDLL(OLIB))
    return _cache


def load_ref_ddms():
    """The reference's own CPU DDMS built from its sources (None if absent)."""
    if not os.path.exists(REFLIB):
        return None
    L = ctypes.CDLL(REFLIB)
    L.ref_ddms_calculate.restype = ctypes.c_float
    L.ref_ddms_calculate.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    return L
