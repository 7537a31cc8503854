// cth_consts.h: constants of the COTHER search hot path shared by the host
// driver and the HIP kernels.  Every value is the one the reference fixes
// at compile time or in its default options; the citation next to each one
// names the reference file (paths relative to the reference's src/).
#pragma once

#include <stdint.h>

// ---- alphabet and profile layout -----------------------------------------
#define CTH_NAA        20   // liblib/alpha.h:10 NUMAA
#define CTH_NCV        19   // PM2DVectorFields.h:44 pmv2DNoCVEls
#define CTH_NSS         3   // liblib/alpha.h:38-43 SS_C, SS_E, SS_H
#define CTH_NTRN        7   // PM2DVectorFields.h:46 (MM,MI,MD,IM,II,DM,DD)
#define CTH_MAXDST     99   // libpro/srcpro/MOptions.h:15 MAX_N_IR_DISTANCE_VALUES
#define CTH_DDSTRIDE  128   // bytes per position of the distance list in HBM (n + 99 values + 0xff pad)
#define CTH_PROSCALE 10000  // libpro/srcpro/PMProfileModel.h:58 ReadWriteScale
#define CTH_MAXCOLUMNS 50000  // liblib/mylimits.h:50 MAXCOLUMNS: longest profile the readers accept
// longest profile whose DP/MAP stripe boundary rows (5 floats per column)
// fit one workgroup's 160 KiB of LDS; longer ones keep them in global memory
#define CTH_LDS_ROW_MAX 8192
// the phase-1 DP keeps the reference's buffer of maxima (value and cell per
// column) next to its five boundary rows: 28 bytes per column in LDS
#define CTH_DP_LDS_ROW_MAX 5824
// the per-query phase-1 DP keeps those rows in global memory from this
// length on: a workgroup of one wavefront with 28 bytes of LDS per column
// leaves most of a CU's SIMDs idle for long profiles
#ifndef CTH_DP_GLOBAL_FROM
#define CTH_DP_GLOBAL_FROM 1536
#endif

// transition order (cudp/CuBatchDP_com.h:87-96)
enum { CTH_MM = 0, CTH_MI, CTH_MD, CTH_IM, CTH_II, CTH_DM, CTH_DD };

// ---- score matrix -----------------------------------------------------------
#define CTH_INITSHIFT     0.002f   // cusco/CuBatchScoreMatrix_Init.cuh:15 (MAPALN=1)
#define CTH_SSSHIFT       0.05f    // cusco/CuBatchScoreMatrix_ssss.cuh:17
#define CTH_CVSSHIFT      0.5f     // cusco/CuBatchScoreMatrix_cvs2s.cuh:18
#define CTH_DDMSSHIFT     0.2f     // cusco/CuBatchScoreMatrix_ddms2s_h2.cuh:20
#define CTH_SSSWGT        0.12f    // env/options.txt SSSWGT
#define CTH_CVSWGT        0.15f    // env/options.txt CVSWGT
#define CTH_DDMSWGT       0.2f     // env/options.txt DDMSWGT
#define CTH_DSTSEGM       6        // env/options.txt DSTSEGM
// distance-match pair score s(d1,d2) for DMS_AS_2 (cupro/SerializedDstMatchScoresAttr.h:32-33,
// cupro/SerializedDstMatchScores.cuh:19-20,350-378)
#define CTH_DMS_S00       0.35f
#define CTH_DMS_S200     (-0.4f)
#define CTH_DMS_MAXDST    20
#define CTH_DMS_MAXDSTINV 0.05f
// DDMS regression (cupro/SerializedDstMatchScoresAttr.h:52,84-91)
#define CTH_DDMS_SCALE    16.0f
#define CTH_DDMS_SLOPE    0.053f
#define CTH_DDMS_ICEPT   (-1.235f)
#define CTH_DDMS_NSEGM    3   // cusco/CuBatchScoreMatrix_ddms2s.cuh:24
#define CTH_DDMS_UNROLL   4   // cusco/CuBatchScoreMatrix_ddms2s_h2.cuh:40
#define CTH_DDMS_STEP     2   // SerializedDstMatchScoresAttr.h:43 DMS_STEP
#define CTH_DDMS_BNDF     2   // SerializedDstMatchScoresAttr.h:51 DMS_DSTBNDF

// ---- phase-1 DP ---------------------------------------------------------------
#define CTH_LOG0          (-32768.0f)  // cudp/CuBatchDP_init.cuh:35-38
#define CTH_OUTSCORE      (-9.0f)      // cudp/CuBatchDP_init_btck.cu:245 (cells outside the valid area)
#define CTH_TRNLOGFCT     0.4f         // libpro/srcpro/TRANSPROBS.h:10
#define CTH_MAXCORNER     60           // cudp/CuBatchDP_com.h:43
#define CTH_CORNERPRC     0.34f        // cudp/CuBatchDP_com.h:46
#define CTH_CORR_NSC      4            // cudp/CuBatchDP_com.h:36
#define CTH_BLK           32           // reference block/diagonal period (cudp/CuBatchDP_com.h:25-27)
// traceback codes (cudp/CuBatchDP_com.h:100-105)
enum { CTH_BT_STOP = 0, CTH_BT_LEFT = 1, CTH_BT_UP = 2, CTH_BT_DIAG = 3 };

// ---- per-profile output record (cuss/CuBatchSS_com.h:45-72) -------------------
enum {
  CTH_O_SCORE = 0, CTH_O_ORGPRONO, CTH_O_PRONEWDST, CTH_O_PSTVS, CTH_O_BEG, CTH_O_END,
  CTH_O_NPHASE1,                       // 6
  CTH_O_LMBDEST = CTH_O_NPHASE1, CTH_O_KEST, CTH_O_LOGE, CTH_O_EPA,
  CTH_O_MIN, CTH_O_MAX, CTH_O_E, CTH_O_H, CTH_O_LMBD, CTH_O_K,
  CTH_O_IDNTS, CTH_O_NGAPS,
  CTH_O_NALNDATA                       // 18
};

// ---- statistics (cuss/CuBatchSS_com.h:16-33) -----------------------------------
#define CTH_SC_MIN       (-10)
#define CTH_SC_MAX         10
#define CTH_NSC            21
#define CTH_SCALE          32
#define CTH_SSC_MIN      (-320)
#define CTH_NSSC          641

// gapped statistics of the scoring system (env/cother.par:14-20)
#define CTH_GAPPED_LAMBDA 0.561132f
#define CTH_GAPPED_K      0.001488f
#define CTH_GAPPED_ALPHA  0.5f
#define CTH_GAPPED_BETA  (-1.0f)
#define CTH_UNGAP_LAMBDA  1.116276f
#define CTH_UNGAP_K       0.115812f
