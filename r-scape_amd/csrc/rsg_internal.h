// rsg_internal.h -- internal state and helpers of librscape_hip (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>
#include <vector>
#include "../../include/rscape_hip.h"

#define RSG_MAX_TIMERS 64

// Pair record of columns i < j (16 cells c[x*4+y]) lives at record index j*L + i, so
// consecutive i (the MFMA result lanes) store to consecutive records.
__host__ __device__ static inline int64_t rsg_rec(int64_t i, int64_t j, int64_t L) { return j * L + i; }

// Kinds of stored pair counts.
//   CNT_I8G: int8-Gram records, cell x*4+y holds count << (x + y) (the MFMA accumulator
//            as produced; readers shift the scale out), 64 bytes per record.
//   CNT_F32: split-BF16 weighted records, FP32 cells, 64 bytes per record.
enum { CNT_I8G = 1, CNT_F32 = 2 };
__host__ __device__ static inline int rsg_cell_shift(int k) { return (k >> 2) + (k & 3); }

struct rsg_timer { const char *name; hipEvent_t a, b; };

struct rsg_ctx {
  int          device = 0;
  hipStream_t  stream = nullptr;
  int          timing = 0;
  int          ntimers = 0;
  rsg_timer    timers[RSG_MAX_TIMERS];
  int          gram_shape = 0;          // 0 = automatic
  // small reusable device scratch (reductions)
  double      *d_red = nullptr;  size_t red_cap = 0;
  int         *d_flag = nullptr;
  double      *h_pinned = nullptr;      // 64 doubles of pinned host staging
};

struct rsg_msa {
  int64_t  N = 0, L = 0;        // sequences, columns
  int64_t  Npad = 0, Lpad = 0;  // padded to the Gram K-tile / column tile
  uint8_t *d_ax  = nullptr;     // N x L raw codes (row-major, as uploaded)
  uint8_t *d_axT = nullptr;     // Lpad x Npad raw codes, transposed (pad = 4, a gap)
  uint8_t *d_xt  = nullptr;     // Lpad x Npad one-hot nibble codes: 1<<c for c<4, else 0
  double  *d_w   = nullptr;     // N weights (FP64); all 1 when unweighted
  int      weighted = 0;
  double   wsum = 0.;
  // split-BF16 weight parts of the weighted Gram: [3][Npad] bf16 bits
  uint16_t *d_wsplit = nullptr;
  int      nsplit_ready = 0;
  // pair tables
  void    *d_cnt = nullptr;     // L x L x 16 (i<j records), kind below
  int      cnt_kind = -1;
  size_t   cnt_bytes = 0;
  void    *d_cnt_u = nullptr;   // unit-weight counts kept for RAF when the main counts are weighted
  int      cnt_u_kind = -1;
  double  *d_pm = nullptr;      // L x 4
  double  *d_ps = nullptr;      // L x 5
  double  *d_lgc = nullptr;     // log table for integer counts, N+2 entries
  double  *d_cov = nullptr;     // L x L (FP64)
  double  *d_tmp = nullptr;     // L x L scratch (RAFS source)
  double  *d_rowx = nullptr;    // L + 1 (row means, then COVavg)
  double   minCOV = 0., maxCOV = 0.;
  int      have_probs = 0, have_cov = 0;
};

// ---- error helpers ----
static inline int rsg_fail(char *errbuf, const char *fmt, ...)
{
  if (errbuf) {
    va_list ap; va_start(ap, fmt);
    vsnprintf(errbuf, RSG_ERRBUFSIZE, fmt, ap);
    va_end(ap);
  }
  return RSG_FAIL;
}

#define HIPCHK(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) \
  return rsg_fail(errbuf, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); } while (0)

// ---- timers ----
static inline void rsg_tstart(rsg_ctx *c, const char *name, int *slot)
{
  *slot = -1;
  if (!c->timing || c->ntimers >= RSG_MAX_TIMERS) return;
  rsg_timer &t = c->timers[c->ntimers];
  if (!t.a) { hipEventCreate(&t.a); hipEventCreate(&t.b); }
  t.name = name;
  hipEventRecord(t.a, c->stream);
  *slot = c->ntimers++;
}
static inline void rsg_tstop(rsg_ctx *c, int slot)
{
  if (slot >= 0) hipEventRecord(c->timers[slot].b, c->stream);
}

// ---- kernel launchers implemented in the .hip files ----
int rsg_pack(rsg_ctx *c, rsg_msa *m, char *errbuf);
int rsg_gram_counts_i8(rsg_ctx *c, rsg_msa *m, void *d_out, int kind, char *errbuf);
int rsg_gram_counts_bf16(rsg_ctx *c, rsg_msa *m, int nsplit, float *d_out, char *errbuf);
int rsg_probs_from_counts(rsg_ctx *c, rsg_msa *m, double tol, char *errbuf);
int rsg_stat(rsg_ctx *c, rsg_msa *m, int base, int cls, const double *allow16, int prec, char *errbuf);
int rsg_raf(rsg_ctx *c, rsg_msa *m, int rafs, const double *allow16, char *errbuf);
int rsg_correct(rsg_ctx *c, rsg_msa *m, int actype, int shiftnonneg, char *errbuf);
int rsg_minmax(rsg_ctx *c, rsg_msa *m, char *errbuf);
int rsg_upload_weights_bf16(rsg_ctx *c, rsg_msa *m, int nsplit, char *errbuf);
int64_t rsg_npad(int64_t N);
