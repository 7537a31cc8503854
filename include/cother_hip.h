/* cother_hip.h: C-ABI of libcother_hip, the MI355X (gfx950) implementation
 * of COTHER's search hot path.
 *
 * The library replaces the device side of the reference's batch-processing
 * interface, class CuBatchProcessing (src/libmycu/cupro/CuBatchProcessing.cuh:
 * 47-244), together with the device-memory manager it drives
 * (src/libmycu/cupro/CuDeviceMemory.cuh:47-770).  Plain pointers and sizes
 * only; every entry point returns 0 on success and a negative code on error,
 * with a message from cth_last_error() (the reference throws
 * myruntime_error through MYCUDACHECK, src/libmycu/cucom/cucommon.h:22-55,
 * and DevCommThread turns it into THREAD_MSG_ERROR,
 * src/libmycu/cualn/DevCommThread.cpp:259-276).
 *
 * Calls on one context are serial (as calls into one CuBatchProcessing
 * instance are, DevCommThread.cpp:296-356); distinct contexts may share a
 * device.
 */
#ifndef COTHER_HIP_H
#define COTHER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- profile sets -----------------------------------------------------------
 * Columnar host profile set; field conventions of the reference's packed
 * profile model (src/libmycu/cupro/PM2DVectorFields.h:53-87) as filled by its
 * profile readers (src/libmycu/cupro/IOProfileModelTxt.cu:296-1030).  The
 * definition is in cother_amd/csrc/cth_pset.h. */
typedef struct cth_pset cth_pset;

/* Allocate an empty set of nprof profiles with the given lengths (positions
 * laid out consecutively).  Replaces the reference's PMBatchProData buffers
 * (src/libmycu/cupro/PMBatchProData.h). */
cth_pset* cth_pset_alloc(int nprof, const int32_t* lens);
void cth_pset_free(cth_pset* ps);

/* Name and description of profile k (NULL when unset): the FILE and DESC
 * fields of a text profile (PMProfileModel.cpp:1893-1894, read at :2377-2384),
 * the description section of the binary database (PMProfileModel.cpp:
 * 1763-1815); the reference keeps them as PMProfileModel::GetName /
 * GetDescription.  cth_pset_set_text copies both (NULL clears). */
const char* cth_pset_name(const cth_pset* ps, int k);
const char* cth_pset_desc(const cth_pset* ps, int k);
int cth_pset_set_text(cth_pset* ps, int k, const char* name, const char* desc);

/* Seeded synthetic profile set in the reference's value conventions (no
 * reference counterpart: the reference has no generator; values follow the
 * text profile round trip of src/libpro/srcpro/PMProfileModel.cpp:1856-2082
 * and the adddist list transform of src/adddist/adddist.cpp:560-651).
 * len_mode 0 uniform [lmin,lmax], 1 lognormal(median,sigma) clipped, 2 fixed;
 * dist_mode 0 predicted maps, 1 C-alpha chains, 2 none; sort_desc orders by
 * length, longest first (makedb order, src/libpro/srcpro/Db.cpp:606). */
cth_pset* cth_synth_make(int nprof, int len_mode, double median, double sigma, int lmin, int lmax,
                         int dist_mode, int sort_desc, uint64_t seed, int nthreads);

/* ---- profile and database files ----------------------------------------------
 * The reference's formats, read into the conventions above.  Text profiles
 * ("COTHER profile v2.5": TextWriteProfileCondensed / TextReadProfile,
 * src/libpro/srcpro/PMProfileModel.cpp:1856-2080, 3527-4620), single files or
 * the makedb text database (.tprd, src/libpro/srcpro/Db.cpp:590-670); the
 * binary database written by db2bin ("COTHER profile bin database v2.5",
 * layout v2.2: Db::MakeBin_v2_2, Db.cpp:337-378, and
 * CalculateAndWriteAddressTable_v2_2 / BinaryWriteProfile_v2_2,
 * PMProfileModel.cpp:1235-1817), which replaces the reference's reader of
 * that file (src/libmycu/cupro/TdDataReader.cpp:385-800).  NULL on error. */
cth_pset* cth_read_text_profiles(const char* path);
cth_pset* cth_read_binary_db(const char* path);
/* Write profiles in the text profile format (TextWriteProfileCondensed);
 * values are scaled integers (scale 10000 in the reference's files). */
int cth_write_text_profiles(const cth_pset* ps, const char* path, int scale);

/* Profile sets from the reference's packed profile-model buffers: pmbeg[k]
 * points at the array of field k of TPM2DVectorFields (181 fields,
 * src/libmycu/cupro/PM2DVectorFields.h:53-87) as TdDataReader::
 * ReadDataChunkHelper packs a database chunk (src/libmycu/cupro/
 * TdDataReader.cpp:703-980) and PMBatchProData the query batch.  nprof < 0:
 * every profile, counted from pmend (each field's end); nprof >= 0: the
 * first nprof profiles from pmbeg (pmend may be NULL), as a worker reads one
 * query once PMDataSkipNPros has advanced the pointers (src/libmycu/cualn/
 * DevCommThread.cpp:314-320).  szfields (optional): each field's byte offset
 * from pmbeg[0] (szCpm2dvfields), checked against pmbeg.  descs (optional):
 * one description per profile.  Field sizes, lengths (1..50000, MAXCOLUMNS)
 * and the position offsets are checked; NULL on error. */
cth_pset* cth_pset_from_packed(char* const* pmbeg, char* const* pmend, const size_t* szfields, int nprof,
                               const char* const* descs);
/* Number of profiles and positions of a set. */
int cth_pset_nprof(const cth_pset* ps);
int64_t cth_pset_npos(const cth_pset* ps);

/* Distance files (adddist: ReadDistances, GetPositionPairHelper,
 * GetPairDistanceHelper and ProcessDistancesForPosition, src/adddist/
 * adddist.cpp:267-651): lines "i j d.. [p..]" of residue pairs (1-based,
 * ordered by i), the distance the mean of the given columns (1-based, >= 3;
 * NULL / 0 columns: column 3), a probability after each distance when
 * prbthrhld > 0.  cth_read_distances writes prolen lists of
 * MAX_N_IR_DISTANCE_VALUES + 1 floats (count, then values);
 * cth_add_distances stores them as profile k's distance lists, rounded as
 * adddist's text output rounds them (PMProfileModel.cpp:2063-2067). */
int cth_read_distances(const char* path, int prolen, const int* columns, int ncolumns, float prbthrhld,
                       float* lists);
int cth_add_distances(cth_pset* ps, int k, const char* path, const int* columns, int ncolumns, float prbthrhld);

/* ---- context ------------------------------------------------------------------ */
typedef struct cth_ctx cth_ctx;

/* Create a context on a device and load the score tables (ssss, cvs2s) from
 * envdir.  Replaces the CuDeviceMemory constructor and its Cache*ScoresDevice
 * calls (src/libmycu/cupro/CuDeviceMemory.cu:100-160,644-830). */
int cth_ctx_create(int device, const char* envdir, cth_ctx** out);
void cth_ctx_destroy(cth_ctx* c);
const char* cth_last_error(void);
/* A context with its own stream and scratch buffers that reads parent's
 * tables, uploaded queries and resident database (one device-memory area of
 * the reference's CuDeviceMemory, src/libmycu/cupro/CuDeviceMemory.cuh:
 * 148-154).  Uploads to it are refused; the parent must outlive it. */
int cth_ctx_create_shared(cth_ctx* parent, cth_ctx** out);
/* Sizes of the uploaded sets (qlens: the context's host copy). */
int cth_ctx_shape(const cth_ctx* c, int* nqueries, const int32_t** qlens, int* ndb, int64_t* dbpos);
/* Number of visible HIP devices (0 when none; the reference's
 * Devices::RegisterDevices, src/libmycu/cualn/Devices.cu:59-108). */
int cth_device_count(void);

/* Upload the query set (CuDeviceMemory::CacheCompleteData query part,
 * src/libmycu/cualn/JobDispatcher.cpp:325-355). */
int cth_upload_queries(cth_ctx* c, const cth_pset* qs);
/* Make a template database resident in HBM (replaces the per-chunk
 * CuBatchProcessing::TransferCPMData, src/libmycu/cupro/CuBatchProcessing.cuh:
 * 60-70, with a one-time upload). */
int cth_upload_db(cth_ctx* c, const cth_pset* db);

/* The two uploads from packed buffers (cth_pset_from_packed): the database
 * chunk as CuBatchProcessing::TransferCPMData receives it
 * (src/libmycu/cupro/CuBatchProcessing.cuh:62-70, CuDeviceMemory::
 * CopyCPMDataToDevice, src/libmycu/cupro/CuDeviceMemory.cu:1095-1157), and
 * the query batch (or its first nprof queries from pmbeg). */
int cth_upload_db_packed(cth_ctx* c, char* const* pmbeg, char* const* pmend, const size_t* szfields);
int cth_upload_queries_packed(cth_ctx* c, char* const* pmbeg, char* const* pmend, const size_t* szfields, int nprof);

/* Score matrix of query qidx against template profiles [p0,p1), copied to
 * host row-major [Lq][ncols] (ncols = positions of those profiles).
 * The stage CuBatchScoreMatrix::ComputeScoreMatrixDevice
 * (src/libmycu/cusco/CuBatchScoreMatrix.cu:122-404) computes on the device;
 * this entry exposes its outputs (scores, modscores) for inspection. */
int cth_score_block(cth_ctx* c, int qidx, int p0, int p1, float* scores, float* modscores);
/* As cth_score_block with the log-odds term from the reduced-precision
 * matrix-core product (contraction 1 BF16, 2 FP16; 0 is cth_score_block). */
int cth_score_block_ex(cth_ctx* c, int qidx, int p0, int p1, int contraction, float* scores, float* modscores);
/* Device check of the matrix-core lane layout the contraction kernel
 * assumes; *bad = number of mismatching accumulator values (0 expected). */
int cth_selftest_mfma(int* bad);

/* Phase 1 of query qidx against template profiles [p0,p1): score matrix,
 * DP with traceback (CuBatchDP::PerformCompleteDynProgDevice,
 * src/libmycu/cudp/CuBatchDP.cu:65-118) and its finalisation (FinalizeDPPro,
 * src/libmycu/cudp/CuBatchDP_final.cu:61-292).  res receives p1-p0 records
 * of six 32-bit words {maxsc(f32), maxxy(u32), passed(i32), score(f32),
 * psts(i32), beg(u32)}; btck (optional) the traceback bytes [Lq][ncols]. */
int cth_phase1_block(cth_ctx* c, int qidx, int p0, int p1, float scorethld, void* res, uint8_t* btck);

/* Phase 1 plus the alignment statistics of the passed profiles
 * (CuBatchSS::CalculateAlnStatisticsDevice, src/libmycu/cuss/CuBatchSS.cu:
 * 63-266): res as cth_phase1_block; plist receives the indices (relative to
 * p0) of the *npass passed profiles; stats npass x 10 doubles {lambda_est,
 * K_est, logE, logEpA, min, max, E, H, lambda, K}. */
int cth_phase2_block(cth_ctx* c, int qidx, int p0, int p1, float scorethld, double searchspace, int ssemodel,
                     void* res, int* plist, int* npass, double* stats);

/* Score matrix of query qidx against template profiles [p0,p1), then the
 * MAP alignment stage (CuBatchMAPDP::PerformCompleteMAPDynProgDevice,
 * src/libmycu/cumapdp/CuBatchMAPDP.cu:67-140) for njobs of those profiles
 * (absolute indices profiles[], log pair E-values logepa[] as phase 2
 * computed them; minpp the MINPP option, default 0.28).  res receives njobs
 * records of nine 32-bit words {maxsc(f32), end(u32), beg(u32), idnts,
 * psts, ngaps, alnlen (i32), logz(f32), maxlog(f32)}. */
int cth_map_block(cth_ctx* c, int qidx, int p0, int p1, const int32_t* profiles, const float* logepa, int njobs,
                  float minpp, void* res);

/* ---- search over the resident database ------------------------------------
 * One call runs the whole per-query pipeline of the reference's
 * CuBatchProcessing::ProcessScoreMatrix (src/libmycu/cupro/
 * CuBatchProcessing.cu:189-575): for each query, score matrix, phase-1 DP,
 * finalisation, statistics of the passed profiles and the E-value filter
 * (JobDispatcher::CalculateScoreThreshold, src/libmycu/cualn/
 * JobDispatcher.cpp:806-860; AbstractScoreMatrix::ComputeLengthAdjustment,
 * src/libpro/srcsco/AbstractScoreMatrix.cpp:247-395), database chunked by
 * mem_budget as the reference chunks by --dev-mem
 * (CuDeviceMemory::CalcMaxDbDataChunkSizeHelper, src/libmycu/cupro/
 * CuDeviceMemory.cu:350-470). */
typedef struct cth_search_opts {
  float eval;         /* E-value threshold (option EVAL, default 10) */
  int ssemodel;       /* statistical significance model (SSEMODEL, default 2) */
  int do_map;         /* realign passed hits by MAP (MAPALN, default 1) */
  int timing;         /* record HIP events around each stage on the context's stream */
  uint64_t mem_budget;/* bytes for the per-chunk score/traceback buffers (0: 64 GiB) */
  float minpp;        /* MAP posterior threshold (MINPP; 0: the default 0.28) */
  int contraction;    /* log-odds term: 0 FP32 as the reference (default); 1 BF16, 2 FP16 operands on
                         the matrix cores with FP32 accumulation (no reference counterpart) */
  char* alnbuf;       /* optional: receives the alignment rows of every hit (see cth_hit.alnoff) */
  uint64_t alncap;    /* capacity of alnbuf in bytes */
  uint64_t dbsize;    /* positions and profiles of the whole database when the resident set is one part of */
  uint64_t ndb;       /* it (CuBatchProcessing::SetDbDetails, src/libmycu/cupro/CuBatchProcessing.cuh:73);
                         both 0: the resident set's own.  They set the score threshold and the search space */
} cth_search_opts;

/* One reported hit; fields follow the reference's per-profile output record
 * (TDP2OutputAlnData, src/libmycu/cuss/CuBatchSS_com.h:45-72). */
typedef struct cth_hit {
  int32_t query, profile;     /* query index, database profile index */
  float score;                /* phase-1 score with the correlation bonus */
  float logE, logEpA;         /* log E-value and log pair E-value */
  float lambda_est, K_est;    /* estimated gapped lambda and K */
  float minsc, maxsc, expected, H, lambda, K;
  int32_t psts;               /* positive substitution scores on the phase-1 path */
  uint32_t beg, end;          /* phase-1 begin/end coordinates (x<<16|y) */
  uint32_t mapbeg, mapend;    /* MAP alignment begin/end coordinates */
  int32_t idnts, ngaps, alnlen, pstvs_map;  /* identities, gaps, length, positives of the MAP alignment */
  float mapscore;             /* MAP alignment score */
  int32_t reserved;
  uint64_t alnoff;            /* with cth_search_opts.alnbuf: offset of this hit's five rows of alnlen
                                 letters each (query, middle, target, query SS, target SS; the output
                                 rows of FinalizeMAPDP_MAP, src/libmycu/cumapdp/CuBatchMAPDP_map_final.cu:
                                 228-300) */
} cth_hit;

typedef struct cth_search_stats {
  double cells;       /* sum over queries of Lq x database positions */
  double ms_phase1;   /* wall time of score + DP + finalisation */
  double ms_phase2;   /* wall time of statistics (+ MAP) */
  double npassed;     /* profiles passed to phase 2 */
  double nchunks;     /* database chunks processed (score-kernel launches) */
  double ms_k_score;  /* with timing: HIP-event time of the score-matrix kernels */
  double ms_k_dp;     /* with timing: HIP-event time of phase-1 DP + finalisation */
  double ms_k_stats;  /* with timing: HIP-event time of the statistics kernels */
  double ms_k_map;    /* with timing: HIP-event time of the MAP kernels */
  double map_cells;   /* sum over realigned hits of Lq x Lt */
} cth_search_stats;

/* Search queries [q0,q1) against the resident database.  hits receives at
 * most maxhits records, grouped by query in query order, each group sorted by
 * (logE, profile); nhits[q-q0] the number of hits of query q; stats (may be
 * NULL) the cells processed and stage times. */
int cth_search(cth_ctx* c, int q0, int q1, const cth_search_opts* opt, cth_hit* hits, int maxhits, int* nhits,
               cth_search_stats* stats);

/* As cth_search, with no device allocation.  No counterpart in the
 * reference: for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Every
 * per-chunk buffer (scores, modular scores, traceback bytes, phase-1
 * records, MAP posteriors, stage lists) is carved out of one fixed scratch
 * array of CTH_NOALLOC_SCRATCH_BYTES that each device holds from the first
 * context created on it until the process ends.  Database chunks take at
 * most three quarters of it (mem_budget is not used); a query whose chunk of
 * a single profile does not fit (Lq x Lt x 13 bytes with MAP, 9 without,
 * about 192 MiB), or a stage whose lists do not fit the remaining quarter,
 * is refused with an error before any kernel of that chunk runs.  The
 * results equal cth_search's.  All calls on one device share the array and
 * it keeps the last call's contents: calls that overlap on one device (from
 * several threads or contexts) race, so the caller serialises them. */
#define CTH_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int cth_search_noalloc(cth_ctx* c, int q0, int q1, const cth_search_opts* opt, cth_hit* hits, int maxhits,
                       int* nhits, cth_search_stats* stats);

/* The same search for many queries at once (no reference counterpart: the
 * reference runs one query per ProcessScoreMatrix call, src/libmycu/cualn/
 * DevCommThread.cpp:310-355, with the host choosing each stage's profiles).
 * Consecutive queries whose buffers fit mem_budget together form a batch
 * that runs each stage once over all its (query, profile) pairs in
 * persistent launches, with the next stage's pairs selected on the device,
 * and one synchronisation per batch; the launches are captured as a HIP
 * graph and replayed while the batch shape stays the same, and the
 * persistent grid sizes are autotuned per batch shape on its first run.
 * Results (hits, their order, alignment rows) equal cth_search's.  Covers
 * the FP32 contraction and databases whose profiles have at most 5824
 * positions; other calls, and queries whose buffers alone exceed the budget,
 * run as cth_search runs them.  With timing, the stages run without the
 * graph and their times fill ms_k_score (score), ms_k_dp (both DP passes),
 * ms_k_stats and ms_k_map; nchunks counts batches. */
int cth_search_batch(cth_ctx* c, int q0, int q1, const cth_search_opts* opt, cth_hit* hits, int maxhits,
                     int* nhits, cth_search_stats* stats);
/* The same for the nq distinct queries qidx[0..nq) in that order (no
 * reference counterpart: the serving loop's ticks batch whichever requests
 * are admitted together); nhits[k] counts query qidx[k]'s hits. */
int cth_search_batch_list(cth_ctx* c, const int32_t* qidx, int nq, const cth_search_opts* opt, cth_hit* hits,
                          int maxhits, int* nhits, cth_search_stats* stats);
/* Graph captures and replays of cth_search_batch on a context. */
int cth_batch_graph_counts(const cth_ctx* c, int64_t* captures, int64_t* replays);

/* Length adjustment and effective search space of a query of qlen positions
 * against dbsize positions in ndb profiles with the gapped parameters
 * (AbstractScoreMatrix::ComputeLengthAdjustment, src/libpro/srcsco/
 * AbstractScoreMatrix.cpp:246-395: GetSearchSpace, GetDeltaLength). */
void cth_search_space(uint64_t qlen, uint64_t dbsize, uint64_t ndb, uint64_t* sspace, uint64_t* deltalen);

/* ---- serving loop -----------------------------------------------------------
 * Requests (one query of the uploaded set each) against the resident
 * database, run concurrently on nslots slots (own stream and scratch per
 * slot) and admitted by buffer footprint: a waiting request starts when a
 * slot is free and the footprints in flight plus its own (Lq x chunk columns
 * x 13 bytes with MAP, 9 without) fit mem_budget; the first waiting request
 * that fits goes first.  Replaces the reference's batch dispatch
 * (JobDispatcher::Run and the DevCommThread workers, src/libmycu/cualn/
 * JobDispatcher.cpp:967-1157, DevCommThread.cpp:149-356) for a resident
 * database.  A request's hits are those cth_search gives for that query. */
typedef struct cth_server cth_server;
typedef struct cth_server_opts {
  int nslots;              /* requests on the device at once */
  uint64_t mem_budget;     /* device bytes for the buffers of the requests in flight */
  uint64_t chunk_cols;     /* database columns per chunk of a request (0: the whole database) */
  cth_search_opts search;  /* per-request options (mem_budget, alnbuf, alncap are set per request) */
  uint64_t alncap;         /* per-request alignment-row capacity (0: no rows) */
  int batch;               /* requests a slot takes in one tick (<= 1: one) */
  uint32_t batch_wait_us;  /* with batch > 1: how long a slot holding fewer than batch requests waits for
                              more to arrive before its tick starts (0: it starts at once) */
} cth_server_opts;
typedef struct cth_request_info {
  int32_t qidx, slot;
  uint64_t footprint;      /* admitted buffer bytes */
  double ms_queued;        /* submit -> start */
  double ms_latency;       /* submit -> done */
  double ms_service;       /* start -> done */
  cth_search_stats stats;  /* of the request's tick (the group's batched search, when batched) */
  int32_t group;           /* requests that ran together in this request's tick */
} cth_request_info;
void cth_server_opts_init(cth_server_opts* o);
/* The parent context holds the uploaded queries and database; create the
 * server after the uploads and destroy it before the parent.  Waiting on
 * every submitted ticket before cth_server_destroy is the caller's part. */
int cth_server_create(cth_ctx* parent, const cth_server_opts* o, cth_server** out);
int cth_server_footprint(const cth_server* s, int qidx, uint64_t* bytes);
int cth_server_submit(cth_server* s, int qidx, uint64_t* ticket);
/* Block until the request is done; hits as cth_search returns them, with
 * alnoff relative to alnbuf when rows were requested. */
int cth_server_wait(cth_server* s, uint64_t ticket, cth_hit* hits, int maxhits, int* nhits, char* alnbuf,
                    uint64_t alncap, cth_request_info* info);
void cth_server_destroy(cth_server* s);

/* ---- results file ----------------------------------------------------------
 * The reference's plain-text output for one query: the per-hit records of
 * CuBatchProcessingFinalizer::CompressResultsPlain (src/libmycu/cupro/
 * CuBatchProcessingFinalizerPlain.cpp:36-403) assembled as
 * AlnWriter::WriteResultsPlain writes them (src/libmycu/cualn/
 * AlnWriterPlain.cpp:19-398, AlnWriter.cpp:250-376). */
typedef struct cth_outinfo {
  const char* qname;          /* query name for the file name (NULL: the query profile's name) */
  const char* qdesc;          /* query description (NULL: the query profile's description) */
  int32_t qrysernr;           /* query serial number (file name suffix __<n>.out) */
  float logevthld;            /* log of the e-value threshold (log EVAL) */
  const char* const* dbnames; /* database names as given on the command line */
  int32_t ndbnames;
  uint64_t dbsize;            /* database positions (also the length adjustment's n) */
  uint64_t ndbentries;        /* database profiles (the length adjustment's N) */
  int32_t nohits, noalns;     /* options NOHITS, NOALNS */
  int32_t dsclen, dscwidth;   /* options DSCLEN, DSCWIDTH */
  int32_t alnwidth, show;     /* options ALNWIDTH, SHOW */
  float ssswgt;               /* option SSSWGT: SS rows are printed when > 0 */
  float ref_lambda, ref_k;    /* ungapped reference values ([UNGAPPED] of cother.par) */
  float gap_lambda, gap_k;    /* gapped reference values */
  const char* progname;       /* first line: program name and version */
  const char* progversion;
} cth_outinfo;

/* Defaults of the reference: options.txt values, cother.par reference
 * values, "cother 2.03.03". */
void cth_outinfo_init(cth_outinfo* o);
/* The hit-list annotation and the alignment record of one hit whose five
 * rows (query, middle, target, query SS, target SS; alnlen letters each) are
 * at rows.  Both strings are allocated; release them with cth_free_text. */
int cth_format_hit_plain(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hit, const char* rows,
                         const cth_outinfo* o, char** annot, char** record);
/* The whole file for query qidx from the hits of a cth_search call with
 * alignment rows (hits of other queries are skipped; reported are those with
 * logE <= logevthld and an alignment, by ascending logE). */
int cth_format_results_plain(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hits, int nhits,
                             const char* alnbuf, uint64_t alncap, const cth_outinfo* o, char** text, uint64_t* len);
/* As cth_format_results_plain, written to <outdir>/<query name>__<qrysernr>.out
 * (AlnWriter::GetOutputFilename); the path is copied to path. */
int cth_write_results_plain(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hits, int nhits,
                            const char* alnbuf, uint64_t alncap, const cth_outinfo* o, const char* outdir, char* path,
                            int pathcap);
/* The same three in the reference's JSON format (CompressResultsJSON,
 * src/libmycu/cupro/CuBatchProcessingFinalizerJSON.cpp:35-383;
 * AlnWriter::WriteResultsJSON, src/libmycu/cualn/AlnWriterJSON.cpp:19-429);
 * the file is <outdir>/<query name>__<qrysernr>.json. */
int cth_format_hit_json(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hit, const char* rows,
                        const cth_outinfo* o, char** annot, char** record);
int cth_format_results_json(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hits, int nhits,
                            const char* alnbuf, uint64_t alncap, const cth_outinfo* o, char** text, uint64_t* len);
int cth_write_results_json(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hits, int nhits,
                           const char* alnbuf, uint64_t alncap, const cth_outinfo* o, const char* outdir, char* path,
                           int pathcap);
void cth_free_text(char* p);

/* Measurement aid (no reference counterpart): a plain device-to-device copy
 * kernel over `bytes` (multiple of 16), launched `reps` times; *ms receives
 * the mean launch time.  Its HBM traffic is known (bytes read + bytes
 * written), which calibrates the traffic counters of the search kernels. */
int cth_bench_copy(int device, size_t bytes, int reps, float* ms);

/* Lane-layout self-checks of the cross-lane moves the DP kernels use (no
 * reference counterpart): shr: out[l] = l-1 for l >= 1, out[0] = -1;
 * shl: out[l] = l+1 for l <= 62, out[63] = -1. */
int cth_selftest_dpp(int* out64);
int cth_selftest_dpp_shl(int* out64);

#ifdef __cplusplus
}
#endif
#endif
