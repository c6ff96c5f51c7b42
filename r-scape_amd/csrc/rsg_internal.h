// rsg_internal.h -- internal state and helpers of librscape_hip (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <cstring>
#include <string>
#include <vector>
#include <thread>
#include <memory>
#include <algorithm>
#include "../../include/rscape_hip.h"

#define RSG_MAX_TIMERS 64

// Pair record of columns i < j (16 cells c[x*4+y]) lives at record index j*L + i, so
// consecutive i (the MFMA result lanes) store to consecutive records.
__host__ __device__ static inline int64_t rsg_rec(int64_t i, int64_t j, int64_t L) { return j * L + i; }

// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror & Shaw, SC'11).  A draw is a
// pure function of (seed, stream, node, position, tag), so simulation is order-free.  One block
// gives four 32-bit words; rsg_philox returns the first.
__host__ __device__ static inline void rsg_philox4(uint64_t seed, uint32_t stream, uint32_t node, uint32_t pos,
                                                   uint32_t tag, uint32_t out[4])
{
  uint32_t c0 = pos, c1 = node, c2 = stream, c3 = tag;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
__host__ __device__ static inline uint32_t rsg_philox(uint64_t seed, uint32_t stream, uint32_t node,
                                                     uint32_t pos, uint32_t tag)
{
  uint32_t w[4];
  rsg_philox4(seed, stream, node, pos, tag, w);
  return w[0];
}
// Tag of the per-node draws of the GTR evolution: the block keyed (internal node v, column)
// gives the left child's symbol draw, the right child's, and the two leaf-gap draws.
#define RSG_TAG_CHILDREN 5u

// Kinds of stored pair counts.
//   CNT_U16: int8-Gram records, 16 uint16 counts (cell x*4+y), 32 bytes per record; for
//            N <= RSG_U16_MAXN sequences.
//   CNT_U32: int8-Gram records for larger N, 16 uint32 counts, 64 bytes per record.
//   CNT_F32: split-BF16 weighted records, FP32 cells, 64 bytes per record.
enum { CNT_U16 = 1, CNT_F32 = 2, CNT_U32 = 3 };
#define RSG_U16_MAXN 65535
static inline int rsg_int_kind(int64_t N) { return N <= RSG_U16_MAXN ? CNT_U16 : CNT_U32; }
__host__ __device__ static inline int rsg_rec_bytes(int kind) { return kind == CNT_U16 ? 32 : 64; }

// One-hot Gram operand ("xt"), K-chunked: [Npad / 16][Lpad][16 bytes], byte = 1 << symbol for
// canonical residues, 0 otherwise.  Writes the four 16-sequence chunk rows of a 64 x 64
// (sequence, column) block held in tile[sequence][column] (256 threads: one 16-byte store
// each, rows contiguous).
__device__ __forceinline__ void pack_xt_chunk(const uint8_t (*tile)[65], int64_t s0, int64_t c0, int64_t Lpad,
                                              uint8_t *__restrict__ xt)
{
  const int q = threadIdx.x >> 6, cc = threadIdx.x & 63;
  unsigned w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int e = 0; e < 16; e++) {
    const unsigned a = tile[q * 16 + e][cc];
    w[e >> 2] |= ((a < 4u) ? (1u << a) : 0u) << (8 * (e & 3));
  }
  *(uint4 *)(xt + ((s0 / 16 + q) * Lpad + c0 + cc) * 16) = make_uint4(w[0], w[1], w[2], w[3]);
}

struct rsg_timer { const char *name; hipEvent_t a, b; double host_ms; };   // host_ms >= 0: a host phase

struct rsg_ctx {
  int          device = 0;
  hipStream_t  stream = nullptr;
  int          timing = 0;
  int          ntimers = 0;
  rsg_timer    timers[RSG_MAX_TIMERS];
  int          gram_shape = 0;          // 0 = automatic
  int          gram_split = 1;          // split the last partial wave of Gram tiles along K
  int          batch_persistent = 1;    // batched kernels: persistent grids (1) or one group per work unit (0)
  int          defer = 0;               // queue status/min/max copies; the caller synchronises once
  int          skip_minmax = 0;         // a correction follows: the statistic's min / max are not needed
  int          ncu = 0;                 // compute units of the device
  void        *d_gscr = nullptr;  size_t gscr_cap = 0;   // Gram split partials
  void        *d_arena = nullptr; size_t arena_cap = 0;  // batched-null buffers (grow-only)
  // the last batched-null graph: instantiated once per launch geometry (sizes, tiles, tree
  // levels, scalar arguments, arena address: the bytes in bexec_sig), replayed by later calls
  // of the same geometry with new inputs
  hipGraphExec_t bexec = nullptr;
  std::vector<uint8_t> bexec_sig;
  int          bexec_hits = 0;          // replays that reused the cached graph (for tests)
  void        *d_hits = nullptr;  size_t hits_cap = 0;   // hit-list compaction (grow-only, entries)
  void        *h_hits = nullptr;                          // its pinned host mirror, same capacity
  // small reusable device scratch (reductions)
  double      *d_red = nullptr;  size_t red_cap = 0;
  int         *d_flag = nullptr;
  // results of one scoring call, in the layout of h_pinned[0..4]: min, max, correction NaN flag
  // (int), COVavg, marginal validation flag (int); one copy brings them to the host, or the
  // chain's last kernel (the min / max reduction) writes them there itself (res_on_host).
  double      *d_res = nullptr;
  double      *h_pinned = nullptr;      // 64 doubles of pinned host staging
  double      *d_hres = nullptr;        // h_pinned as the device addresses it
  int          res_on_host = 0;
  // completion stamp: the min / max reduction writes hstamp to h_pinned[63] after the result
  // block (system-scope release), so the caller can wait for it without a stream synchronisation
  unsigned long long hstamp = 0;
  int          stamp_wait_us = 2000;    // spin on the stamp this long, then synchronise the stream
  void        *d_small = nullptr;       // 64 KB of device scratch for small per-call tables
  // set for the duration of an allocation-free call (rsg_noalloc.hip): the Gram takes this
  // 32 x 32-column tile list instead of its cached one, and d_red is not grown past red_cap
  const int2  *tiles_fixed = nullptr;
  int          ntiles_fixed = 0;
  int          no_alloc = 0;
};

struct rsg_msa {
  int64_t  N = 0, L = 0;        // sequences, columns
  int64_t  Npad = 0, Lpad = 0;  // padded to the Gram K-tile / column tile
  uint8_t *d_ax  = nullptr;     // N x L raw codes (row-major, as uploaded)
  uint8_t *d_axT = nullptr;     // Lpad x Npad raw codes, transposed (pad = 4, a gap)
  uint8_t *d_xt  = nullptr;     // Lpad x Npad one-hot nibble codes: 1<<c for c<4, else 0
  double  *d_w   = nullptr;     // N weights (FP64); all 1 when unweighted
  int      weighted = 0;
  int      weights_ok = 1;    // every weight finite and >= 0 (else pp cannot validate)
  double   wsum = 0.;
  // split-BF16 weight parts of the weighted Gram: [3][Npad] bf16 bits
  uint16_t *d_wsplit = nullptr;
  int      nsplit_ready = 0;
  uint8_t *d_wq = nullptr;      // int8 weight limbs [Npad/16][4][16] and (after them) the FP64 scale
  int      wq_ready = 0;
  int      wq_counts = 0;       // the current counts come from the rounded weights (RSG_COUNTS_W8)
  double  *d_nsx = nullptr;     // with wq_counts: nseff of every pair (record index) in FP64
  double   wsum_q = 0.;         // sum of the rounded weights W * q_s
  // pair tables
  void    *d_cnt = nullptr;     // L x L x 16 (i<j records), kind below
  int      cnt_kind = -1;
  int      pm_fused = 0;          // the last int8 Gram left the marginal partials in d_red
  size_t   cnt_bytes = 0;
  void    *d_cnt_u = nullptr;   // unit-weight counts kept for RAF when the main counts are weighted
  int      cnt_u_kind = -1;
  double  *d_pm = nullptr;      // L x 4
  double  *d_ps = nullptr;      // L x 5
  double  *d_lgc = nullptr;     // log table for integer counts, N+2 entries
  int64_t  lgc_n = -1;          // the N the log table holds (rebuilt when it differs)
  double  *d_cov = nullptr;     // L x L (FP64)
  double  *d_tmp = nullptr;     // L x L scratch (RAFS source)
  double  *d_rowx = nullptr;    // 2 L + 2: row means, row sums, COVavg
  double  *d_eval = nullptr;    // L x L E-values (kept on the device for the hit list)
  int      have_eval = 0;
  double   minCOV = 0., maxCOV = 0.;
  int      have_probs = 0, have_cov = 0;
};

// Total weight the counts were formed with (ngap = total - nseff): the rounded weights' sum for
// counts from int8 weight limbs, so ngap is the gapped sequences' rounded weight
static inline double rsg_count_wsum(const rsg_msa *m) { return m->wq_counts ? m->wsum_q : m->wsum; }


// Host threads for planning work: the hardware's, at most 16 (one GPU's share of a node).
static inline int rsg_host_threads()
{
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::min(16u, std::max(1u, h));
}

// fn(k0, k1) over contiguous ranges of [0, n) of about equal total weight w(k) (total: their
// sum), one range per host thread (fewer when the total is under min_work per thread); the
// calling thread takes the first range.
template <class W, class F>
static void rsg_parallel_ranges(int n, int64_t total, W w, F fn, int64_t min_work = 1 << 16)
{
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)rsg_host_threads(), (int64_t)n, total / min_work + 1}));
  std::vector<int> rb((size_t)nt + 1, n);
  rb[0] = 0;
  int64_t acc = 0;
  for (int k = 0, t = 1; k < n && t < nt; k++) {
    acc += w(k);
    while (t < nt && acc * nt >= total * t) rb[(size_t)t++] = k + 1;
  }
  std::vector<std::thread> th;
  for (int t = 1; t < nt; t++) th.emplace_back(fn, rb[(size_t)t], rb[(size_t)t + 1]);
  fn(rb[0], rb[1]);
  for (auto &x : th) x.join();
}

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
  if (c->timing == 2 && strncmp(name, "gram", 4) != 0) return;     // Gram launches only
  rsg_timer &t = c->timers[c->ntimers];
  if (!t.a) { hipEventCreate(&t.a); hipEventCreate(&t.b); }
  t.name = name; t.host_ms = -1.;
  hipEventRecord(t.a, c->stream);
  *slot = c->ntimers++;
}
static inline void rsg_tstop(rsg_ctx *c, int slot)
{
  if (slot >= 0) hipEventRecord(c->timers[slot].b, c->stream);
}
// a host-side phase measured by the caller (wall clock, ms), reported with the device timers
static inline void rsg_thost(rsg_ctx *c, const char *name, double ms)
{
  if (c->timing != 1 || c->ntimers >= RSG_MAX_TIMERS) return;
  rsg_timer &t = c->timers[c->ntimers++];
  t.name = name; t.host_ms = ms;
}

// ---- kernel launchers implemented in the .hip files ----
int rsg_noalloc_init(int device, char *errbuf);   // rsg_noalloc.hip: the scratch array's address
int rsg_pack(rsg_ctx *c, rsg_msa *m, char *errbuf);
int rsg_msa_alloc(rsg_ctx *c, int64_t nseq, int64_t alen, rsg_msa **ret, char *errbuf);   // rsg_capi.hip
// pm_part (uint16 records only; nullptr = none): marginal partials of every 32 x 32 block
// (pm32 layout, rsg_device.h); pmflag: cleared for the marginal reduction that follows
int rsg_gram_counts_i8(rsg_ctx *c, rsg_msa *m, void *d_out, int kind, double *pm_part, int *pmflag, char *errbuf);
int rsg_gram_counts_bf16(rsg_ctx *c, rsg_msa *m, int nsplit, float *d_out, char *errbuf);
int rsg_gram_counts_w8(rsg_ctx *c, rsg_msa *m, float *d_out, char *errbuf);
int rsg_probs_from_counts(rsg_ctx *c, rsg_msa *m, double tol, char *errbuf);
double *rsg_pm32_scratch(rsg_ctx *c, int64_t L, char *errbuf);
int rsg_stat(rsg_ctx *c, rsg_msa *m, int base, int cls, const double *allow16, int prec, char *errbuf);
int rsg_raf(rsg_ctx *c, rsg_msa *m, int rafs, const double *allow16, char *errbuf);
int rsg_correct(rsg_ctx *c, rsg_msa *m, int actype, int shiftnonneg, char *errbuf);
int rsg_reserve_scratch(rsg_ctx *c, int64_t L, char *errbuf);
int rsg_minmax(rsg_ctx *c, rsg_msa *m, char *errbuf);
int rsg_upload_weights_bf16(rsg_ctx *c, rsg_msa *m, int nsplit, char *errbuf);
int64_t rsg_npad(int64_t N);

// One simulated alignment: tree + output rows (rsg_sim.hip).
struct SimItem {
  int64_t N, L;           // leaves, columns
  int64_t tree_off;       // first internal node of this item in the concatenated tree arrays
  int64_t seq_off;        // offset of its (N-1) x L internal-node symbol rows
  int64_t leaf_off;       // offset of its N x L leaf output
  uint64_t seed; uint32_t stream; uint32_t gapthr;   // a leaf symbol is gapped if its tag-1 draw < gapthr
  int64_t col_blk0;       // first 64-column block of this item in the launch
  int64_t model_off;      // offset (doubles) of its 40-double GTR model
};
// Tree levels of a set of items: entries {item, internal node, first block within the level, 0}
// grouped by depth.
// Host staging array without value-initialisation (every element is written before it is read).
template <class T> struct HostArr {
  std::unique_ptr<T[]> p;
  size_t n = 0;
  void alloc(size_t k) { p.reset(new T[k]); n = k; }
  T *data() { return p.get(); }
  const T *data() const { return p.get(); }
  size_t size() const { return n; }
  T &operator[](size_t i) { return p[i]; }
  const T &operator[](size_t i) const { return p[i]; }
};

struct SimLevels {
  HostArr<int4> ent;
  std::vector<int64_t> lev_ent0;     // entry offset of each level (size nlev + 1)
  std::vector<int64_t> lev_blocks;   // 64-column blocks of each level
  std::vector<int64_t> lev_maxl;     // longest alignment among each level's entries
};
int rsg_sim_levels(const SimItem *items, int nitems, const int *left, const int *right, SimLevels &lv, char *errbuf);
int rsg_sim_launch(rsg_ctx *c, const SimItem *d_items, int nitems, int64_t nodes, int64_t blocks,
                   const SimLevels &lv, const int4 *d_ent, const int *d_left, const int *d_right,
                   const double *d_ld, const double *d_rd, const double *d_models, uint32_t *d_cdf,
                   uint8_t *d_seqs, uint8_t *d_leaves, const uint8_t *root0, int64_t root_len);
int rsg_simulate_device(rsg_ctx *c, int nitems, const int64_t *N, const int64_t *L,
                        const int *left, const int *right, const double *ld, const double *rd,
                        const double *models, const uint64_t *seed, const uint32_t *stream, const double *gapfrac,
                        uint8_t *d_leaves, const int64_t *leaf_off, char *errbuf, const uint8_t *root0 = nullptr);
// one alignment's leaves to the host, the root drawn from pi (root == nullptr) or given (L codes)
int rsg_sim_one(rsg_ctx *c, const rsg_null_job *job, const uint8_t *root, uint8_t *leaves, char *errbuf);
int rsg_sim_pairs(rsg_ctx *c, const rsg_null_job *job, const int *ct, const double *model16, const uint8_t *root,
                  uint8_t *leaves, char *errbuf);

// R-scape's null of one alignment for R replicates at once (rsg_fitch.hip): the tree's levels
// (internal nodes grouped by depth, nodes[lev0[k] .. lev0[k+1]) at depth k), parents and the
// widest level, then one launch per kernel and depth covering every replicate.
struct FitchPlan {
  int64_t N = 0, L = 0;
  int m = 0, maxd = 0, maxw = 0;
  std::vector<int> parent, nodes, lev0;
};
// Device buffers of R replicates: ax (N x L codes), the tree (left, right, parent, nodes: m
// each), perm (R x L), seeds (R), mask and all (R x (2N - 1) x L), inner (R x (N - 1) x L:
// the shuffled internal rows), leaves (R x N x L: the null alignments), nsub (R x 2m x 25),
// cols (R x 2 maxw x L ints).
struct FitchDev {
  const uint8_t *ax;
  const int *left, *right, *parent, *nodes, *perm;
  const uint64_t *seeds;
  uint8_t *mask, *all, *inner, *leaves;
  int *nsub, *cols;
};
int rsg_fitch_plan(int64_t N, int64_t L, const int *left, const int *right, FitchPlan &P, char *errbuf);
void rsg_fitch_perm(uint64_t seed, int64_t L, int *perm);
int rsg_fitch_launch(rsg_ctx *c, const FitchPlan &P, const FitchDev &D, int R);
int rsg_fitch_ancestral(rsg_ctx *c, const FitchPlan &P, const FitchDev &D, int R);

// Batched pipeline: per-item descriptor of RSG_BDESC int64 fields
// {N, L, Npad, Lpad, off_xt, off_cnt, off_cov, off_pm, off_rows, (unused), off_leaf, blk0_pack,
//  blk0_col, blk0_stat, blk0_el, nb32, group, blk0_hist, off_pm32 (marginal partials),
//  blk0_col16 (16-column blocks)}
#define RSG_BDESC 20
// Surrogate features (rsg_features.hip): column statistics of items {one-hot panel offset, N,
// Npad, L, first 4-column block} into 36 uint64 counters per item; the host turns them into
// RSG_NULL_NFEAT floats.
int rsg_feat_launch(rsg_ctx *c, const int64_t *d_items, int nitems, int64_t nblocks, const uint8_t *d_base,
                    unsigned long long *d_out);
void rsg_feat_finish(const unsigned long long *cnt, int64_t N, int64_t L, const int *left, const int *right,
                     const double *ld, const double *rd, float *f);
// Large alignments of a batch use 64 x 32-column tiles (4 waves, 256-sequence K tiles), the
// others 32 x 32-column tiles (2 waves, 64-sequence K tiles).
static inline bool rsg_batch_big_tiles(int64_t N, int64_t L) { return L >= 256 && N >= 256; }
// d_queue: 128 ints of device memory for the persistent kernel's per-XCD tile counters (unused
// by the one-work-group-per-tile launch, rsg_ctx::batch_persistent = 0).
int rsg_gram_batch_launch(rsg_ctx *c, const int4 *d_big, int nbig, const int4 *d_small, int nsmall,
                          const int64_t *d_desc, const uint8_t *d_xt, uint16_t *d_cnt, double *d_pm_part, int pm_field,
                          int *d_queue);
