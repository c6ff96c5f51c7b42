// rsg_noalloc.hip -- the allocation-free scoring entry point (rsg_cov_Calculate_noalloc).
//
// For callers that cannot allocate device memory (code under stream capture, tight serving
// loops, harnesses that forbid hipMalloc): the whole cov_Calculate chain of one alignment of up
// to RSG_NA_MAXN sequences x RSG_NA_MAXL columns runs with its intermediate buffers (transposed
// codes, one-hot panel, pair records, marginals, log table, row sums, reduction partials, Gram
// tile list) carved from `scratch`, one file-scope device array of fixed size per GPU, held until
// the process ends.  The caller passes the input codes and the output matrix as device buffers.
// Every call on a GPU shares that GPU's array and leaves its contents behind; each call rebuilds
// everything it reads, so no state carries over, but calls that overlap on one GPU (two host
// threads, two contexts) race on it.
#include "rsg_internal.h"
#include <mutex>

#define NA_ALIGN 256
static constexpr size_t na_round(size_t b) { return (b + NA_ALIGN - 1) / NA_ALIGN * NA_ALIGN; }
// Largest padded shapes: Npad is a multiple of 256 (RSG_NA_MAXN is one), Lpad of 64.
static constexpr size_t NA_AXT   = na_round((size_t)RSG_NA_MAXL * RSG_NA_MAXN);          // transposed codes
static constexpr size_t NA_XT    = NA_AXT;                                                // one-hot panel
static constexpr size_t NA_CNT   = na_round((size_t)RSG_NA_MAXL * RSG_NA_MAXL * 32);    // uint16 records
static constexpr size_t NA_PM    = na_round(sizeof(double) * 8 * RSG_NA_MAXL);
static constexpr size_t NA_PS    = na_round(sizeof(double) * 5 * RSG_NA_MAXL);
static constexpr size_t NA_LGC   = na_round(sizeof(double) * 3 * (RSG_NA_MAXN + 1));
static constexpr size_t NA_ROWX  = na_round(sizeof(double) * (2 * RSG_NA_MAXL + 2));
static constexpr size_t NA_REDN  = 196608;                                                // doubles
static constexpr size_t NA_RED   = na_round(sizeof(double) * NA_REDN);
static constexpr size_t NA_TILES = na_round(sizeof(int2) * (RSG_NA_MAXL / 32) * (RSG_NA_MAXL / 32 + 1) / 2);
static constexpr size_t NA_BYTES = NA_AXT + NA_XT + NA_CNT + NA_PM + NA_PS + NA_LGC + NA_ROWX + NA_RED + NA_TILES;

// The scratch array.  In host code its name is a host-side stand-in, not a device address: the
// device address is taken once per device with hipGetSymbolAddress (rsg_noalloc_init) and only
// that address is used.
static __device__ __align__(256) uint8_t scratch[NA_BYTES];

#define NA_MAXDEV 64
static void *g_scratch[NA_MAXDEV];
static std::mutex g_scratch_mu;

// Called by rsg_ctx_create with the context's device current: the device address of this
// device's copy of the array (once per device).
int rsg_noalloc_init(int device, char *errbuf)
{
  if (device < 0 || device >= NA_MAXDEV) return rsg_fail(errbuf, "device %d beyond the %d the library tracks", device, NA_MAXDEV);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return RSG_OK;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return rsg_fail(errbuf, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  g_scratch[device] = ptr;
  return RSG_OK;
}

// The 32 x 32-column tiles of the pair triangle in the Gram's order (bi ascending, then
// bj = bi .. nb - 1); the last diagonal tile is left out when its block holds one column.
__global__ void k_tiles32(int nb, int ntiles, int2 *__restrict__ out)
{
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= ntiles) return;
  int bi = 0;
  while ((bi + 1) * nb - ((bi + 1) * bi) / 2 <= t) bi++;        // first entry of row bi + 1 > t
  const int off = bi * nb - (bi * (bi - 1)) / 2;
  out[t] = make_int2(bi, bi + (t - off));
}

extern "C" int rsg_cov_Calculate_noalloc(rsg_ctx *c, const uint8_t *d_ax, int64_t nseq, int64_t alen, int covtype,
                                         int covclass, int prec, const double *allowpair, int nseqthresh,
                                         int alenthresh, double tol, double *d_cov, double *minCOV, double *maxCOV,
                                         char *errbuf)
{
  if (!c || !d_ax || !d_cov) return rsg_fail(errbuf, "rsg_cov_Calculate_noalloc: null context or buffer");
  if (nseq < 1 || alen < 1) return rsg_fail(errbuf, "empty alignment (nseq %lld, alen %lld)", (long long)nseq, (long long)alen);
  if (nseq > RSG_NA_MAXN || alen > RSG_NA_MAXL)
    return rsg_fail(errbuf, "rsg_cov_Calculate_noalloc: %lld x %lld exceeds the fixed scratch (at most %d sequences x %d columns)",
                    (long long)nseq, (long long)alen, RSG_NA_MAXN, RSG_NA_MAXL);
  const int base = covtype - covtype % 3;
  if (covtype < 0 || base == RSG_RAF || base == RSG_RAFS || base == RSG_CCF || base > RSG_CCF)
    return rsg_fail(errbuf, "rsg_cov_Calculate_noalloc: covariation type %d not available (CHI/GT/MI/MIr/MIg/OMES only)", covtype);
  if (c->device < 0 || c->device >= NA_MAXDEV || !g_scratch[c->device])
    return rsg_fail(errbuf, "rsg_cov_Calculate_noalloc: no scratch address for device %d", c->device);
  HIPCHK(hipSetDevice(c->device));
  uint8_t *s = (uint8_t *)g_scratch[c->device];
  // the alignment as the chain sees it: every buffer from the scratch, the output the caller's
  rsg_msa m;
  m.N = nseq; m.L = alen;
  m.Npad = rsg_npad(nseq);
  m.Lpad = ((alen + 63) / 64) * 64;
  m.d_ax = (uint8_t *)d_ax;
  m.d_axT = s;                        s += NA_AXT;
  m.d_xt = s;                         s += NA_XT;
  m.d_cnt = s;                        s += NA_CNT;
  m.d_pm = (double *)s;               s += NA_PM;
  m.d_ps = (double *)s;               s += NA_PS;
  m.d_lgc = (double *)s;              s += NA_LGC;
  m.d_rowx = (double *)s;             s += NA_ROWX;
  double *red = (double *)s;          s += NA_RED;
  int2 *tiles = (int2 *)s;
  m.d_cov = d_cov;
  m.cnt_kind = CNT_U16;
  m.cnt_bytes = (size_t)32 * (size_t)alen * (size_t)alen;
  m.weighted = 0;
  m.wsum = (double)nseq;
  m.lgc_n = -1;                       // the table is rebuilt for this N
  const int nb = (int)((alen + 31) / 32);
  const int ntiles = nb * (nb + 1) / 2 - ((alen - 32 * (int64_t)(nb - 1)) == 1 ? 1 : 0);
  // the context runs this call with fixed resources and restores its own afterwards
  const int sv_shape = c->gram_shape, sv_split = c->gram_split, sv_timing = c->timing;
  double *sv_red = c->d_red; const size_t sv_cap = c->red_cap;
  c->gram_shape = 4;                  // 32 x 32-column tiles, 2 waves, 128-sequence K tiles
  c->gram_split = 0;                  // no split-K partials
  c->timing = 0;                      // no timer events
  c->d_red = red; c->red_cap = NA_REDN;
  c->tiles_fixed = tiles; c->ntiles_fixed = ntiles;
  c->no_alloc = 1;
  int st = RSG_OK;
  if (ntiles > 0) {
    hipLaunchKernelGGL(k_tiles32, dim3((unsigned)((ntiles + 255) / 256)), dim3(256), 0, c->stream, nb, ntiles, tiles);
    if (hipGetLastError() != hipSuccess) st = rsg_fail(errbuf, "rsg_cov_Calculate_noalloc: tile list launch failed");
  }
  if (st == RSG_OK) st = rsg_pack(c, &m, errbuf);
  if (st == RSG_OK)
    st = rsg_cov_Calculate(c, &m, covtype, covclass, RSG_COUNTS_INT8, prec, allowpair, nseqthresh, alenthresh, tol,
                           minCOV, maxCOV, errbuf);
  c->gram_shape = sv_shape; c->gram_split = sv_split; c->timing = sv_timing;
  c->d_red = sv_red; c->red_cap = sv_cap;
  c->tiles_fixed = nullptr; c->ntiles_fixed = 0;
  c->no_alloc = 0;
  return st;
}
