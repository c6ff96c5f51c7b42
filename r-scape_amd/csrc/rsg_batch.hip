// rsg_batch.hip -- batched phylogenetic null: simulate, score and fit many alignments of
// uneven size in one pass on the device.
//
// Per alignment this is R-scape's per-replicate work (null_rscape, src/R-scape.c:1615-1724
// with the cov_GenerateAlignment simulator, src/cov_simulate.c:60-190): cov_Calculate of the
// replicate (corr_Probs + statistic + APC/ASC, src/covariation.c:63-306), the histogram width
// of calculate_width_histo (src/R-scape.c:1280-1371), the rank histogram
// (src/covariation.c:414-432) and the tail fit (cov_NullFitExponential / Gamma,
// :1914-1975 with cov_histogram_pmass :2483-2506).
//
// Layout: every alignment ("item") owns slices of a few large buffers, described by an
// int64 descriptor (RSG_BDESC fields, see rsg_internal.h).  Each kernel below covers all
// items in one launch; a block finds its item by binary search over a per-kernel block
// prefix.  No host synchronisation happens between the launches, so the whole batch can
// be recorded once as a hipGraph and replayed; buffers are cleared by fill kernels, not
// memset nodes.
#include "rsg_device.h"
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cstring>
#include <numeric>

// Histogram geometry, bins and E-values are rounded exactly as the host computes them.
#pragma clang fp contract(off)

enum { D_N = 0, D_L, D_NPAD, D_LPAD, D_XT, D_CNT, D_COV, D_PM, D_ROWS, D_HIST, D_LEAF, D_BPACK, D_BCOL,
       D_BSTAT, D_BEL, D_NB32, D_GROUP, D_BHIST, D_PMP, D_BCOL16 };
#define HIST_EPB 65536       // matrix elements per histogram block
#define NBLDS 16384          // histograms up to this many bins are privatised in LDS per block
#define NBPOOL 16384         // bins of the shared bin pool per group (a group may take more)
#define ELPB 4096            // matrix elements per block in the element-wise kernels

__device__ __forceinline__ int find_item(const int64_t *__restrict__ desc, int n, int field, int64_t blk)
{
  int lo = 0, hi = n - 1;
  while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (desc[mid * RSG_BDESC + field] <= blk) lo = mid; else hi = mid - 1; }
  return lo;
}

// leaves (N x L, row-major) -> one-hot codes, K-chunked (pack_xt_chunk)
__global__ void __launch_bounds__(256)
k_pack_b(const int64_t *__restrict__ desc, int n, const uint8_t *__restrict__ leaves, uint8_t *__restrict__ xt)
{
  __shared__ uint8_t tile[64][65];
  const int it = find_item(desc, n, D_BPACK, blockIdx.x);
  const int64_t *d = desc + it * RSG_BDESC;
  const int64_t N = d[D_N], L = d[D_L], Lpad = d[D_LPAD];
  const int64_t local = blockIdx.x - d[D_BPACK], nbx = Lpad / 64;
  const int64_t c0 = (local % nbx) * 64, s0 = (local / nbx) * 64;
  const uint8_t *ax = leaves + d[D_LEAF];
  uint8_t *X = xt + d[D_XT];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < 64; r += 4) {
    int64_t s = s0 + r, c = c0 + tx;
    tile[r][tx] = (s < N && c < L) ? ax[s * L + c] : (uint8_t)4;
  }
  __syncthreads();
  pack_xt_chunk(tile, s0, c0, Lpad, X);
}

// corr_Marginals per item from the 32 x 32 block partials the batched Gram leaves (the
// single-alignment path's layout, order and reduction: a replicate's marginals match its
// rsg_corr_Probs bit for bit)
__global__ void __launch_bounds__(256)
k_pm_reduce_b(const int64_t *__restrict__ desc, int n, const double *__restrict__ part, double *__restrict__ pm,
              double tol, int *__restrict__ flag)
{
  const int it = find_item(desc, n, D_BCOL16, blockIdx.x);
  const int64_t *d = desc + it * RSG_BDESC;
  const int64_t L = d[D_L], nb32 = d[D_NB32];
  const int64_t col = (blockIdx.x - d[D_BCOL16]) * 16 + (threadIdx.x >> 4);
  if (col >= L) return;
  const double *pi = part + d[D_PMP];
  pm32_reduce_col(pi, pi + pm32_n(nb32), (int)nb32, L, col, threadIdx.x & 15, tol, pm + d[D_PM], pm + d[D_PM] + 4 * L, flag);
}

__global__ void __launch_bounds__(256)
k_stat_b(const int64_t *__restrict__ desc, int n, StatArgs A0, const uint16_t *__restrict__ cnt,
         const double *__restrict__ pm, double *__restrict__ cov, const int *__restrict__ cls, int base)
{
  const int it = find_item(desc, n, D_BSTAT, blockIdx.x);
  const int64_t *d = desc + it * RSG_BDESC;
  const int64_t L = d[D_L], local = blockIdx.x - d[D_BSTAT];
  StatArgs A = A0;
  A.cnt = cnt + d[D_CNT]; A.pm = pm + d[D_PM]; A.lpm = pm + d[D_PM] + 4 * L; A.L = L;
  A.wsum = (double)d[D_N]; A.cov = cov + d[D_COV]; A.type = base; A.cls = cls[it];
  int bi, bj;
  tri_tile(local, bi, bj);                      // tiles on or above the diagonal only
  stat_tile<double, CNT_U16, true>(A, bi, bj, (int64_t)blockIdx.x);
}

// raw row sums of COV (diagonal excluded): one wave per row, four rows per block.  Lane l
// keeps the partials of the single path's threads l, l + 64, l + 128, l + 192 (k_rowsum's
// 256-strided sums) and combines them in block_sum_tree256's order, so every row sum equals
// the single path's bit for bit.
__global__ void __launch_bounds__(256)
k_rowsum_b(const int64_t *__restrict__ desc, int n, int64_t nrows, const double *__restrict__ cov, double *__restrict__ rows)
{
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= nrows) return;
  const int lane = threadIdx.x & 63;
  const int it = find_item(desc, n, D_BCOL, row);
  const int64_t *d = desc + it * RSG_BDESC;
  const int64_t L = d[D_L], i = row - d[D_BCOL];
  const double *C = cov + d[D_COV] + i * L;
  double p[4] = {0., 0., 0., 0.};
#pragma unroll
  for (int q = 0; q < 4; q++)
    for (int64_t j = lane + 64 * q; j < L; j += 256) if (j != i) p[q] += C[j];
  double x = (p[0] + p[2]) + (p[1] + p[3]);
#pragma unroll
  for (int st = 32; st > 0; st >>= 1) x += __shfl_down(x, st, 64);
  if (lane == 0) rows[d[D_ROWS] + L + i] = x;
}

__global__ void __launch_bounds__(256)
k_rowmean_b(const int64_t *__restrict__ desc, double *__restrict__ rows)
{
  __shared__ double red[256];
  const int64_t *d = desc + blockIdx.x * RSG_BDESC;
  const int64_t L = d[D_L];
  double *R = rows + d[D_ROWS];
  double s = 0.;
  for (int64_t i = threadIdx.x; i < L; i += 256) { double r = R[L + i]; s += r; R[i] = (L > 1) ? r / ((double)L - 1.) : r; }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) { if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st]; __syncthreads(); }
  if (threadIdx.x == 0) R[2 * L] = (L > 1) ? red[0] / ((double)L * ((double)L - 1.)) : red[0];
}

// APC / ASC in place on both triangles; block min / max to part[blockIdx]
__global__ void __launch_bounds__(256)
k_correct_b(const int64_t *__restrict__ desc, int n, double *__restrict__ cov, const double *__restrict__ rows,
            int actype, int *__restrict__ flag, double *__restrict__ part)
{
  __shared__ double rmin[256], rmax[256];
  const int it = find_item(desc, n, D_BEL, blockIdx.x);
  const int64_t *d = desc + it * RSG_BDESC;
  const int64_t L = d[D_L], e0 = (blockIdx.x - d[D_BEL]) * ELPB;
  double *C = cov + d[D_COV];
  const double *R = rows + d[D_ROWS];
  const double avg = R[2 * L];
  double a = INFINITY, b = -INFINITY;
  const int64_t e1 = min(e0 + (int64_t)ELPB, L * L);
  int64_t e = e0 + threadIdx.x;
  int64_t i = e / L, j = e - i * L;                 // then stepped by 256 without a division
  for (; e < e1; e += 256) {
    if (i != j) {
      double v = C[e];
      double r = (actype == RSG_APC) ? ((avg != 0.0) ? v - R[i] * R[j] / avg : 0.0) : v - (R[i] + R[j] - avg);
      if (isnan(r)) atomicOr(flag, 1);
      C[e] = r;
      a = fmin(a, r); b = fmax(b, r);
    }
    j += 256;
    while (j >= L) { j -= L; i++; }
  }
  rmin[threadIdx.x] = a; rmax[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) { rmin[threadIdx.x] = fmin(rmin[threadIdx.x], rmin[threadIdx.x + s]); rmax[threadIdx.x] = fmax(rmax[threadIdx.x], rmax[threadIdx.x + s]); }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = rmin[0]; part[2 * blockIdx.x + 1] = rmax[0]; }
}

// per item: reduce block partials [p0, p1) to (min, max)
__global__ void __launch_bounds__(256)
k_item_minmax(const int64_t *__restrict__ desc, int field, int64_t nblk_total, int n, const double *__restrict__ part,
              double *__restrict__ mm)
{
  __shared__ double rmin[256], rmax[256];
  const int it = blockIdx.x;
  const int64_t p0 = desc[it * RSG_BDESC + field];
  const int64_t p1 = (it + 1 < n) ? desc[(it + 1) * RSG_BDESC + field] : nblk_total;
  double a = INFINITY, b = -INFINITY;
  for (int64_t k = p0 + threadIdx.x; k < p1; k += 256) { a = fmin(a, part[2 * k]); b = fmax(b, part[2 * k + 1]); }
  rmin[threadIdx.x] = a; rmax[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) { rmin[threadIdx.x] = fmin(rmin[threadIdx.x], rmin[threadIdx.x + s]); rmax[threadIdx.x] = fmax(rmax[threadIdx.x], rmax[threadIdx.x + s]); }
    __syncthreads();
  }
  if (threadIdx.x == 0) { mm[2 * it] = rmin[0]; mm[2 * it + 1] = rmax[0]; }
}

// Histogram geometry per null (group of replicates): the width comes from the group's first
// replicate (calculate_width_histo runs on the first null replicate, src/R-scape.c:1353);
// the bounds cover every replicate, as the cumulative histogram grows to do
// (null_add2cumranklist, src/R-scape.c:1564-1612).
struct HGeo { double w, bmax, xmin, xmax, npairs, mn, mx; int64_t off; int nb, imin, imax, ok; };

__global__ void k_width_b(int ng, const int *__restrict__ gfirst, const int64_t *__restrict__ desc,
                          const double *__restrict__ mm, double bmin, double w0, int hpts, double tol,
                          HGeo *__restrict__ geo)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const int j0 = gfirst[g], j1 = gfirst[g + 1];
  HGeo G; G.ok = 0;
  const double lo = fmax(bmin, mm[2 * j0]);
  double w = fmin(w0, (mm[2 * j0 + 1] - lo) / (double)hpts);
  if (w < tol) w = 0.;
  double mn = INFINITY, mx = -INFINITY, np = 0.;
  for (int j = j0; j < j1; j++) {
    mn = fmin(mn, mm[2 * j]); mx = fmax(mx, mm[2 * j + 1]);
    const int64_t L = desc[j * RSG_BDESC + D_L];
    np += (double)(L * (L - 1) / 2);
  }
  G.w = w; G.bmax = 0.; G.xmin = G.xmax = 0.; G.nb = 0; G.imin = 0; G.imax = -1; G.off = 0;
  G.npairs = np; G.mn = mn; G.mx = mx;
  if (w > 0. && mx > bmin) {
    double bmax = mx + 5 * w;
    while (fabs(bmax - bmin) < tol) bmax += w;
    const double nbd = (bmax - bmin) / w;
    if (nbd < (double)INT_MAX) {
      G.nb = (int)nbd; G.bmax = bmax;
      G.xmin = fmax(mn, bmin + w); G.xmax = fmax(mx, bmin + w);
      G.imin = (int)ceil((G.xmin - bmin) / w - 1.); G.imax = (int)ceil((G.xmax - bmin) / w - 1.);
      G.ok = (G.imin >= 0 && G.imax < G.nb) ? 1 : 0;
    }
  }
  geo[g] = G;
}

// Bins of every group in one pool: offsets by an exclusive scan over the groups (one block,
// fixed order); a group that does not fit in the pool's cap bins is marked failed (its fit is
// NaN, as for a group without a histogram).
__global__ void __launch_bounds__(1024) k_hist_layout(int ng, HGeo *__restrict__ geo, int64_t cap)
{
  __shared__ int64_t sh[1024];
  const int T = 1024, tid = threadIdx.x;
  const int chunk = (ng + T - 1) / T, g0 = tid * chunk, g1 = min(g0 + chunk, ng);
  int64_t own = 0;
  for (int g = g0; g < g1; g++) own += geo[g].ok ? (int64_t)geo[g].nb : 0;
  sh[tid] = own;
  __syncthreads();
  for (int o = 1; o < T; o <<= 1) {                 // inclusive Hillis-Steele scan
    const int64_t add = (tid >= o) ? sh[tid - o] : 0;
    __syncthreads();
    sh[tid] += add;
    __syncthreads();
  }
  int64_t off = sh[tid] - own;
  for (int g = g0; g < g1; g++) {
    if (!geo[g].ok) continue;
    geo[g].off = off;
    off += geo[g].nb;
    if (off > cap) geo[g].ok = 0;
  }
}

// clear the bins of every group (blocks take groups round-robin)
__global__ void __launch_bounds__(256) k_clear_bins(int ng, const HGeo *__restrict__ geo, unsigned long long *__restrict__ hist)
{
  for (int g = blockIdx.x; g < ng; g += gridDim.x) {
    const HGeo G = geo[g];
    if (!G.ok) continue;
    for (int64_t b = threadIdx.x; b < G.nb; b += 256) hist[G.off + b] = 0ull;
  }
}

__global__ void k_fill_int(int *__restrict__ p, int n)
{
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0;
}

// Rank histogram of every item into its group's bins: each block bins up to HIST_EPB
// matrix elements into an LDS histogram (32-bit, at most HIST_EPB per bin), then adds its
// nonzero bins to the group histogram.
__global__ void __launch_bounds__(256)
k_hist_b(const int64_t *__restrict__ desc, int n, const double *__restrict__ cov, const HGeo *__restrict__ geo,
         double bmin, unsigned long long *__restrict__ hist)
{
  __shared__ unsigned int lh[NBLDS];
  const int it = find_item(desc, n, D_BHIST, blockIdx.x);
  const int64_t *d = desc + it * RSG_BDESC;
  const HGeo g = geo[d[D_GROUP]];
  if (!g.ok) return;
  const bool lds = (g.nb <= NBLDS);                 // else straight to the group's bins
  unsigned long long *H = hist + g.off;
  if (lds) for (int b = threadIdx.x; b < g.nb; b += 256) lh[b] = 0u;
  __syncthreads();
  const int64_t L = d[D_L], e0 = (blockIdx.x - d[D_BHIST]) * (int64_t)HIST_EPB;
  const int64_t e1 = min(e0 + (int64_t)HIST_EPB, L * L);
  const double *C = cov + d[D_COV];
  const double fl = bmin + g.w;
  int64_t e = e0 + threadIdx.x;
  int64_t i = e / L, j = e - i * L;                 // then stepped by 256 without a division
  for (; e < e1; e += 256) {
    if (i < j) {
      const double x = C[e];
      const double add = (x > fl) ? x : fl;
      int b = (int)ceil(((add - bmin) / g.w) - 1.);
      b = (b < 0) ? 0 : ((b >= g.nb) ? g.nb - 1 : b);
      if (lds) atomicAdd(&lh[b], 1u);
      else     atomicAdd(&H[b], 1ull);
    }
    j += 256;
    while (j >= L) { j -= L; i++; }
  }
  __syncthreads();
  if (lds)
    for (int b = threadIdx.x; b < g.nb; b += 256)
      if (lh[b]) atomicAdd(&H[b], (unsigned long long)lh[b]);
}

// tail fit per null (one 1024-thread block per group): see rsg_rank.hip for the method
__global__ void __launch_bounds__(1024)
k_fit_b(const HGeo *__restrict__ geo, const unsigned long long *__restrict__ hist,
        unsigned long long *__restrict__ S, int *__restrict__ NZ, double bmin, double pmass, double fracfit, int doexp,
        double *__restrict__ fits)
{
  const int g = blockIdx.x;
  const HGeo G = geo[g];
  double *o = fits + g * 8;
  if (threadIdx.x == 0) { o[6] = G.mn; o[7] = G.mx; }
  if (!G.ok) {
    if (threadIdx.x == 0) { o[0] = NAN; o[1] = INFINITY; o[2] = 0.; o[3] = 0.; o[4] = NAN; o[5] = G.w; }
    return;
  }
  nullfit_body(hist + G.off, G.nb, G.imin, G.imax, G.npairs, G.npairs, bmin, G.w, pmass, fracfit, doexp,
               S + G.off, NZ + G.off, nullptr, o);
  if (threadIdx.x == 0) { o[5] = G.w; o[6] = G.mn; o[7] = G.mx; }
}

// the single-alignment table (k_logtab) for counts up to NM
__global__ void k_logtab_b(int64_t NM, double *__restrict__ t)
{
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k <= NM) {
    const double l = log((double)k + 1e-10);
    t[k] = l; t[NM + 1 + k] = log((double)k + 16e-10); t[2 * (NM + 1) + k] = ((double)k + 1e-10) * l;
  }
}

// ------------------------------------------------------------------------------- host side
namespace {
// Bump allocation over a base pointer (256-byte aligned slices); with base == nullptr it only
// measures.
struct Arena {
  uint8_t *base = nullptr;
  size_t off = 0;
  template <typename T> T *get(size_t count)
  {
    T *p = base ? (T *)(base + off) : nullptr;
    off += (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~(size_t)255;
    return p;
  }
};
}

// The bytes that fix a batch's launch geometry, kept to compare a call with the cached graph's
struct GeoSig {
  std::vector<uint8_t> b;
  void add(const void *p, size_t n)
  {
    const uint64_t len = n;
    b.insert(b.end(), (const uint8_t *)&len, (const uint8_t *)&len + sizeof(len));
    if (n) b.insert(b.end(), (const uint8_t *)p, (const uint8_t *)p + n);
  }
};

// The batch; with fp != nullptr the null alignments are R-scape's nulls of one alignment fax
// (Fitch + shuffle, rsg_fitch.hip), one per job (all jobs of that alignment's size, seeds from the
// jobs), in place of the simulation.
static int null_batch_impl(rsg_ctx *c, const rsg_null_job *jobs, int nj, int covtype, double bmin, double pmass,
                           double fracfit, int doexpfit, rsg_hist **hists, double *fits, float *feats, int use_graph,
                           const uint8_t *fax, const FitchPlan *fp, char *errbuf)
{
  HIPCHK(hipSetDevice(c->device));
  if (nj < 1) return rsg_fail(errbuf, "empty batch");
  const int base = covtype - covtype % 3, corr = covtype % 3;
  auto clk = std::chrono::steady_clock::now;
  auto since = [](std::chrono::steady_clock::time_point a) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
  };
  auto h0 = clk();
  // groups: contiguous, ascending from 0
  std::vector<int> gfirst;
  for (int k = 0; k < nj; k++) {
    const int g = jobs[k].group;
    if (k == 0 ? g != 0 : (g != jobs[k - 1].group && g != jobs[k - 1].group + 1))
      return rsg_fail(errbuf, "job %d: group ids must start at 0 and be contiguous and ascending", k);
    if (k == 0 || g != jobs[k - 1].group) gfirst.push_back(k);
  }
  const int ng = (int)gfirst.size();
  gfirst.push_back(nj);
  if (base == RSG_RAF || base == RSG_RAFS || base == RSG_CCF)
    return rsg_fail(errbuf, "batched null supports CHI/GT/MI/MIr/MIg/OMES statistics");
  // ---- layout
  std::vector<int64_t> desc((size_t)nj * RSG_BDESC, 0);
  std::vector<SimItem> sim((size_t)nj);
  std::vector<int> cls((size_t)nj);
  int64_t o_xt = 0, o_cnt = 0, o_cov = 0, o_pm = 0, o_rows = 0, o_leaf = 0, nodes = 0, seqs = 0;
  int64_t b_pack = 0, b_col = 0, b_stat = 0, b_el = 0, b_sim = 0, b_hist = 0, nmax = 1;
  int64_t b_col16 = 0, o_pmp = 0;
  std::vector<int4> tiles, btiles;        // 32 x 32-column tiles (small alignments), 64 x 32 (large)
  for (int k = 0; k < nj; k++) {
    const rsg_null_job &J = jobs[k];
    if (J.nleaf < 2 || J.alen < 2) return rsg_fail(errbuf, "job %d: needs >= 2 sequences and >= 2 columns", k);
    if (J.nleaf > RSG_U16_MAXN) return rsg_fail(errbuf, "job %d: more than %d sequences", k, RSG_U16_MAXN);
    const int64_t N = J.nleaf, L = J.alen, Npad = rsg_npad(N), Lpad = ((L + 63) / 64) * 64;
    int64_t *d = &desc[(size_t)k * RSG_BDESC];
    d[D_N] = N; d[D_L] = L; d[D_NPAD] = Npad; d[D_LPAD] = Lpad;
    d[D_XT] = o_xt; d[D_CNT] = o_cnt; d[D_COV] = o_cov; d[D_PM] = o_pm; d[D_ROWS] = o_rows;
    d[D_HIST] = 0; d[D_LEAF] = o_leaf;
    d[D_BPACK] = b_pack; d[D_BCOL] = b_col; d[D_BSTAT] = b_stat; d[D_BEL] = b_el;
    const int64_t nb32 = (L + 31) / 32;
    d[D_NB32] = nb32; d[D_GROUP] = J.group; d[D_BHIST] = b_hist;
    d[D_PMP] = o_pmp; d[D_BCOL16] = b_col16;
    b_col16 += (L + 15) / 16; o_pmp += 2 * pm32_n(nb32);
    cls[(size_t)k] = (N <= 8 || L <= 50) ? RSG_C2 : RSG_C16;           // CSELECT (correlators.c:335-341)
    SimItem &s = sim[(size_t)k];
    s.N = N; s.L = L; s.tree_off = nodes; s.seq_off = seqs; s.leaf_off = o_leaf; s.seed = J.seed; s.stream = J.stream;
    s.gapthr = (J.gapfrac <= 0.) ? 0u : (J.gapfrac >= 1. ? 0xFFFFFFFFu : (uint32_t)(J.gapfrac * 4294967296.0));
    s.col_blk0 = b_sim; s.model_off = 40 * (int64_t)k;
    o_xt += Lpad * Npad; o_cnt += 16 * L * L; o_cov += L * L; o_pm += 8 * L; o_rows += 2 * L + 2; o_leaf += N * L;
    nodes += N - 1; seqs += (N - 1) * L;
    b_pack += (Lpad / 64) * (Npad / 64); b_col += L; b_stat += nb32 * (nb32 + 1) / 2; b_el += (L * L + ELPB - 1) / ELPB;
    b_sim += (L + 63) / 64; b_hist += (L * L + HIST_EPB - 1) / HIST_EPB;
    nmax = std::max(nmax, N);
  }
  rsg_thost(c, "batch_plan_layout", since(h0));
  h0 = clk();
  // Gram tiles, longest K loops first so the tail of each launch is short work: the jobs in
  // order of decreasing padded sequence count (ties in job order), each job's tiles row by row
  {
    std::vector<int> ord((size_t)nj);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
      return desc[(size_t)a * RSG_BDESC + D_NPAD] > desc[(size_t)b * RSG_BDESC + D_NPAD];
    });
    for (int k : ord) {
      const int64_t N = jobs[k].nleaf, L = jobs[k].alen, Lpad = desc[(size_t)k * RSG_BDESC + D_LPAD], nb32 = (L + 31) / 32;
      if (rsg_batch_big_tiles(N, L)) {
        // 64 (i) x 32 (j) columns: every tile holding a pair i < j < L
        for (int64_t bi = 0; bi < Lpad / 64; bi++)
          for (int64_t bj = 0; bj < nb32; bj++)
            if (bi * 64 < std::min((bj + 1) * 32, L) - 1) btiles.push_back(make_int4(k, (int)bi, (int)bj, 0));
      } else {
        for (int64_t bi = 0; bi < nb32; bi++)
          for (int64_t bj = bi; bj < nb32; bj++) tiles.push_back(make_int4(k, (int)bi, (int)bj, 0));
      }
    }
  }
  rsg_thost(c, "batch_plan_tiles", since(h0));
  h0 = clk();
  HostArr<int> left, right;
  HostArr<double> ld, rd, models;
  SimLevels lv;
  std::vector<int> fperm, ftree;
  std::vector<uint64_t> fseed;
  if (!fp) {
    left.alloc((size_t)nodes); right.alloc((size_t)nodes);
    ld.alloc((size_t)nodes); rd.alloc((size_t)nodes); models.alloc((size_t)nj * 40);
    rsg_parallel_ranges(nj, nodes, [&](int k) { return jobs[k].nleaf - 1; }, [&](int k0, int k1) {
      for (int k = k0; k < k1; k++) {
        const int64_t off = sim[(size_t)k].tree_off, m1 = jobs[k].nleaf - 1;
        memcpy(&left[(size_t)off], jobs[k].left, sizeof(int) * m1);
        memcpy(&right[(size_t)off], jobs[k].right, sizeof(int) * m1);
        memcpy(&ld[(size_t)off], jobs[k].ld, sizeof(double) * m1);
        memcpy(&rd[(size_t)off], jobs[k].rd, sizeof(double) * m1);
        memcpy(&models[(size_t)k * 40], jobs[k].model, sizeof(double) * 40);
      }
    });
    rsg_thost(c, "batch_plan_trees", since(h0));
    h0 = clk();
    if (rsg_sim_levels(sim.data(), nj, left.data(), right.data(), lv, errbuf) != RSG_OK) return RSG_FAIL;
    rsg_thost(c, "batch_plan_levels", since(h0));
    h0 = clk();
  } else {
    // one column permutation per replicate; the tree as {left, right, parent, level-ordered nodes}
    for (int k = 0; k < nj; k++)
      if (jobs[k].nleaf != fp->N || jobs[k].alen != fp->L) return rsg_fail(errbuf, "Fitch batch: job %d has another size", k);
    fperm.resize((size_t)nj * fp->L);
    fseed.resize((size_t)nj);
    for (int k = 0; k < nj; k++) { fseed[(size_t)k] = jobs[k].seed; rsg_fitch_perm(jobs[k].seed, fp->L, &fperm[(size_t)k * fp->L]); }
    ftree.resize(4 * (size_t)fp->m);
    memcpy(&ftree[0], jobs[0].left, sizeof(int) * fp->m);
    memcpy(&ftree[(size_t)fp->m], jobs[0].right, sizeof(int) * fp->m);
    memcpy(&ftree[2 * (size_t)fp->m], fp->parent.data(), sizeof(int) * fp->m);
    memcpy(&ftree[3 * (size_t)fp->m], fp->nodes.data(), sizeof(int) * fp->m);
  }
  // surrogate features of each group's first replicate
  std::vector<int64_t> fitems((size_t)ng * 6);
  int64_t fblocks = 0;
  for (int g = 0; g < ng; g++) {
    const int64_t *d0 = &desc[(size_t)gfirst[(size_t)g] * RSG_BDESC];
    int64_t *f = &fitems[(size_t)g * 6];
    f[0] = d0[D_XT]; f[1] = d0[D_N]; f[2] = d0[D_NPAD]; f[3] = d0[D_L]; f[4] = fblocks; f[5] = d0[D_LPAD];
    fblocks += (d0[D_L] + 3) / 4;
  }
  // histogram bins of all groups share one pool (a group's share is set on the device once its
  // score range is known)
  const int64_t pool = std::max<int64_t>((int64_t)ng * NBPOOL, (int64_t)1 << 21);
  // ---- device buffers: one grow-only arena of the context, carved in two passes
  int64_t *d_desc; SimItem *d_sim; int *d_cls; int4 *d_tiles, *d_btiles; int *d_left, *d_right;
  double *d_ld, *d_rd, *d_models; int4 *d_ent; int *d_gfirst; int64_t *d_fitems; unsigned long long *d_fcnt;
  uint32_t *d_cdf; uint8_t *d_seqs, *d_leaves, *d_xt; uint16_t *d_cnt; double *d_cov, *d_pm, *d_rows, *d_lgc;
  double *d_part_s, *d_part_e, *d_mm, *d_pmp; HGeo *d_geo; unsigned long long *d_hist, *d_S; int *d_NZ; double *d_fits;
  uint8_t *f_ax = nullptr, *f_mask = nullptr, *f_all = nullptr; int *f_tree = nullptr, *f_perm = nullptr, *f_nsub = nullptr;
  int *f_cols = nullptr; uint64_t *f_seeds = nullptr; int *d_queue = nullptr;
  auto carve = [&](Arena &A) {
    d_desc = A.get<int64_t>(desc.size()); d_sim = A.get<SimItem>((size_t)nj); d_cls = A.get<int>((size_t)nj);
    d_tiles = A.get<int4>(tiles.size()); d_btiles = A.get<int4>(btiles.size());
    d_left = A.get<int>((size_t)nodes); d_right = A.get<int>((size_t)nodes);
    d_ld = A.get<double>((size_t)nodes); d_rd = A.get<double>((size_t)nodes);
    d_models = A.get<double>(40 * (size_t)nj); d_ent = A.get<int4>(lv.ent.size());
    d_gfirst = A.get<int>((size_t)ng + 1); d_fitems = A.get<int64_t>(fitems.size());
    d_fcnt = A.get<unsigned long long>((size_t)ng * 36);
    d_cdf = A.get<uint32_t>((size_t)nodes * 24); d_seqs = A.get<uint8_t>((size_t)seqs);
    d_leaves = A.get<uint8_t>((size_t)o_leaf); d_xt = A.get<uint8_t>((size_t)o_xt);
    d_cnt = A.get<uint16_t>((size_t)o_cnt); d_cov = A.get<double>((size_t)o_cov); d_pm = A.get<double>((size_t)o_pm);
    d_rows = A.get<double>((size_t)o_rows); d_lgc = A.get<double>(3 * (size_t)(nmax + 1));
    d_part_s = A.get<double>(2 * (size_t)b_stat); d_part_e = A.get<double>(2 * (size_t)b_el);
    d_mm = A.get<double>(2 * (size_t)nj); d_pmp = A.get<double>((size_t)o_pmp); d_geo = A.get<HGeo>((size_t)ng);
    d_hist = A.get<unsigned long long>((size_t)pool); d_S = A.get<unsigned long long>((size_t)pool);
    d_NZ = A.get<int>((size_t)pool); d_fits = A.get<double>((size_t)ng * 8); d_queue = A.get<int>(128);
    if (fp) {
      const size_t N = (size_t)fp->N, L = (size_t)fp->L, R = (size_t)nj;
      f_ax = A.get<uint8_t>(N * L); f_tree = A.get<int>(4 * (size_t)fp->m); f_perm = A.get<int>(R * L);
      f_seeds = A.get<uint64_t>(R); f_mask = A.get<uint8_t>(R * (2 * N - 1) * L); f_all = A.get<uint8_t>(R * (2 * N - 1) * L);
      f_nsub = A.get<int>(R * 2 * (size_t)fp->m * 25); f_cols = A.get<int>(R * 2 * (size_t)fp->maxw * L);
    }
  };
  rsg_thost(c, "batch_plan_rest", since(h0));
  h0 = clk();
  size_t plan_bytes = 0;
  {
    Arena plan;
    carve(plan);
    plan_bytes = plan.off;
    if (c->arena_cap < plan.off) {
      if (c->bexec) { hipGraphExecDestroy(c->bexec); c->bexec = nullptr; c->bexec_sig.clear(); }
      if (c->d_arena) hipFree(c->d_arena);
      c->d_arena = nullptr; c->arena_cap = 0;
      if (hipMalloc(&c->d_arena, plan.off) != hipSuccess)
        return rsg_fail(errbuf, "batch: out of device memory (%zu bytes)", plan.off);
      c->arena_cap = plan.off;
    }
    Arena A; A.base = (uint8_t *)c->d_arena;
    carve(A);
  }
  HIPCHK(hipMemcpyAsync(d_desc, desc.data(), sizeof(int64_t) * desc.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_sim, sim.data(), sizeof(SimItem) * nj, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_cls, cls.data(), sizeof(int) * nj, hipMemcpyHostToDevice, c->stream));
  if (!tiles.empty()) HIPCHK(hipMemcpyAsync(d_tiles, tiles.data(), sizeof(int4) * tiles.size(), hipMemcpyHostToDevice, c->stream));
  if (!btiles.empty()) HIPCHK(hipMemcpyAsync(d_btiles, btiles.data(), sizeof(int4) * btiles.size(), hipMemcpyHostToDevice, c->stream));
  if (!fp) {
    HIPCHK(hipMemcpyAsync(d_left, left.data(), sizeof(int) * nodes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d_right, right.data(), sizeof(int) * nodes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d_ld, ld.data(), sizeof(double) * nodes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d_rd, rd.data(), sizeof(double) * nodes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d_models, models.data(), sizeof(double) * models.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d_ent, lv.ent.data(), sizeof(int4) * lv.ent.size(), hipMemcpyHostToDevice, c->stream));
  } else {
    HIPCHK(hipMemcpyAsync(f_ax, fax, (size_t)(fp->N * fp->L), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(f_tree, ftree.data(), sizeof(int) * ftree.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(f_perm, fperm.data(), sizeof(int) * fperm.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(f_seeds, fseed.data(), sizeof(uint64_t) * fseed.size(), hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(hipMemcpyAsync(d_gfirst, gfirst.data(), sizeof(int) * gfirst.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_fitems, fitems.data(), sizeof(int64_t) * fitems.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  rsg_thost(c, "batch_host_upload", since(h0));
  h0 = clk();

  StatArgs A0;
  memset(&A0, 0, sizeof(A0));
  A0.lgc = d_lgc; A0.N = nmax; A0.part = d_part_s;
  for (int k = 0; k < 16; k++) {                     // WC + GU (src/R-scape.c:882-887)
    int x = k / 4, y = k % 4;
    if ((x == 0 && y == 3) || (x == 3 && y == 0) || (x == 1 && y == 2) || (x == 2 && y == 1) ||
        (x == 2 && y == 3) || (x == 3 && y == 2)) A0.allow |= 1u << k;
  }

  auto launch_all = [&]() -> int {
    if (fp) {
      const int m = fp->m;
      const FitchDev D{f_ax, f_tree, f_tree + m, f_tree + 2 * m, f_tree + 3 * m, f_perm, f_seeds, f_mask, f_all, d_seqs,
                       d_leaves, f_nsub, f_cols};
      if (rsg_fitch_launch(c, *fp, D, nj) != RSG_OK) return RSG_FAIL;
    } else if (rsg_sim_launch(c, d_sim, nj, nodes, b_sim, lv, d_ent, d_left, d_right, d_ld, d_rd, d_models, d_cdf, d_seqs,
                              d_leaves, nullptr, 0) != RSG_OK) return RSG_FAIL;
    hipLaunchKernelGGL(k_pack_b, dim3((unsigned)b_pack), dim3(256), 0, c->stream, d_desc, nj, d_leaves, d_xt);
    if (feats && rsg_feat_launch(c, d_fitems, ng, fblocks, d_xt, d_fcnt) != RSG_OK) return RSG_FAIL;
    if (rsg_gram_batch_launch(c, d_btiles, (int)btiles.size(), d_tiles, (int)tiles.size(), d_desc, d_xt, d_cnt, d_pmp,
                              (int)D_PMP, d_queue) != RSG_OK)
      return RSG_FAIL;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, c->d_flag, 16);
    hipLaunchKernelGGL(k_logtab_b, dim3((unsigned)((nmax + 256) / 256)), dim3(256), 0, c->stream, nmax, d_lgc);
    hipLaunchKernelGGL(k_pm_reduce_b, dim3((unsigned)b_col16), dim3(256), 0, c->stream, d_desc, nj, d_pmp, d_pm, 1e-6, c->d_flag);
    hipLaunchKernelGGL(k_stat_b, dim3((unsigned)b_stat), dim3(256), 0, c->stream, d_desc, nj, A0, d_cnt, d_pm, d_cov, d_cls, base);
    if (corr) {
      hipLaunchKernelGGL(k_rowsum_b, dim3((unsigned)((b_col + 3) / 4)), dim3(256), 0, c->stream, d_desc, nj, b_col, d_cov, d_rows);
      hipLaunchKernelGGL(k_rowmean_b, dim3((unsigned)nj), dim3(256), 0, c->stream, d_desc, d_rows);
      hipLaunchKernelGGL(k_correct_b, dim3((unsigned)b_el), dim3(256), 0, c->stream, d_desc, nj, d_cov, d_rows,
                         corr == 1 ? RSG_APC : RSG_ASC, c->d_flag + 1, d_part_e);
      hipLaunchKernelGGL(k_item_minmax, dim3((unsigned)nj), dim3(256), 0, c->stream, d_desc, (int)D_BEL, b_el, nj, d_part_e, d_mm);
    } else {
      hipLaunchKernelGGL(k_item_minmax, dim3((unsigned)nj), dim3(256), 0, c->stream, d_desc, (int)D_BSTAT, b_stat, nj, d_part_s, d_mm);
    }
    hipLaunchKernelGGL(k_width_b, dim3((unsigned)((ng + 63) / 64)), dim3(64), 0, c->stream, ng, d_gfirst, d_desc, d_mm,
                       bmin, 0.05, 400, 1e-6, d_geo);
    hipLaunchKernelGGL(k_hist_layout, dim3(1), dim3(1024), 0, c->stream, ng, d_geo, pool);
    hipLaunchKernelGGL(k_clear_bins, dim3((unsigned)std::min(ng, 4096)), dim3(256), 0, c->stream, ng, d_geo, d_hist);
    hipLaunchKernelGGL(k_hist_b, dim3((unsigned)b_hist), dim3(256), 0, c->stream, d_desc, nj, d_cov, d_geo, bmin, d_hist);
    hipLaunchKernelGGL(k_fit_b, dim3((unsigned)ng), dim3(1024), 0, c->stream, d_geo, d_hist, d_S, d_NZ,
                       bmin, pmass, fracfit, doexpfit, d_fits);
    return hipGetLastError() == hipSuccess ? RSG_OK : RSG_FAIL;
  };

  int t; rsg_tstart(c, "null_batch", &t);
  if (use_graph) {
    // The graph's nodes depend only on the launch geometry: the descriptors (sizes and offsets),
    // the tile lists, the tree levels, the scalar arguments and the arena's address.  Every
    // input that changes between calls of one geometry (seeds, trees, branch lengths, models,
    // the alignment, permutations) is read from device buffers uploaded above, outside the
    // graph, so a graph instantiated for this key is replayed as it is.
    GeoSig H;
    H.add(&nj, sizeof(nj)); H.add(&ng, sizeof(ng)); H.add(&covtype, sizeof(covtype)); H.add(&bmin, sizeof(bmin));
    H.add(&pmass, sizeof(pmass)); H.add(&fracfit, sizeof(fracfit)); H.add(&doexpfit, sizeof(doexpfit));
    const int flags2[3] = {feats != nullptr, fp != nullptr, c->batch_persistent};
    H.add(flags2, sizeof(flags2));
    H.add(desc.data(), sizeof(int64_t) * desc.size());
    H.add(tiles.data(), sizeof(int4) * tiles.size()); H.add(btiles.data(), sizeof(int4) * btiles.size());
    const int64_t sc[6] = {nodes, b_sim, fblocks, pool, nmax, (int64_t)plan_bytes};
    H.add(sc, sizeof(sc));
    H.add(lv.lev_ent0.data(), sizeof(int64_t) * lv.lev_ent0.size());
    H.add(lv.lev_blocks.data(), sizeof(int64_t) * lv.lev_blocks.size());
    H.add(lv.lev_maxl.data(), sizeof(int64_t) * lv.lev_maxl.size());
    if (fp) {
      const int64_t fs[5] = {fp->N, fp->L, fp->m, fp->maxd, fp->maxw};
      H.add(fs, sizeof(fs)); H.add(fp->lev0.data(), sizeof(int) * fp->lev0.size());
    }
    const uintptr_t base_addr = (uintptr_t)c->d_arena;
    H.add(&base_addr, sizeof(base_addr));
    if (!c->bexec || c->bexec_sig != H.b) {
      if (c->bexec) { hipGraphExecDestroy(c->bexec); c->bexec = nullptr; c->bexec_sig.clear(); }
      hipGraph_t graph = nullptr;
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int lr = launch_all();
      hipError_t ec = hipStreamEndCapture(c->stream, &graph);
      if (lr != RSG_OK || ec != hipSuccess) { if (graph) hipGraphDestroy(graph); return rsg_fail(errbuf, "batch graph capture failed"); }
      hipGraphExec_t exec = nullptr;
      hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (ei != hipSuccess) return rsg_fail(errbuf, "graph instantiate failed");
      c->bexec = exec; c->bexec_sig.swap(H.b);
    } else c->bexec_hits++;
    hipError_t el = hipSuccess, es = hipSuccess;
    for (int rep = 0; rep < use_graph && el == hipSuccess && es == hipSuccess; rep++) {
      el = hipGraphLaunch(c->bexec, c->stream);             // every replay recomputes from scratch
      if (rep + 1 < use_graph) es = hipStreamSynchronize(c->stream);
    }
    rsg_tstop(c, t);
    if (es == hipSuccess) es = hipStreamSynchronize(c->stream);
    if (el != hipSuccess || es != hipSuccess) {
      hipGraphExecDestroy(c->bexec); c->bexec = nullptr; c->bexec_sig.clear();
      return rsg_fail(errbuf, "batch graph launch failed");
    }
  } else {
    if (launch_all() != RSG_OK) return rsg_fail(errbuf, "batch launch failed");
    rsg_tstop(c, t);
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  rsg_thost(c, "batch_host_run", since(h0));
  int flags[2];
  HIPCHK(hipMemcpy(flags, c->d_flag, sizeof(flags), hipMemcpyDeviceToHost));
  if (flags[0]) return rsg_fail(errbuf, "pm validation failed in a batch item");
  if (flags[1]) return rsg_fail(errbuf, "bad covariation (NaN after correction) in a batch item");
  HIPCHK(hipMemcpy(fits, d_fits, sizeof(double) * 8 * ng, hipMemcpyDeviceToHost));
  if (feats) {
    std::vector<unsigned long long> cnt((size_t)ng * 36);
    HIPCHK(hipMemcpy(cnt.data(), d_fcnt, sizeof(unsigned long long) * cnt.size(), hipMemcpyDeviceToHost));
    for (int g = 0; g < ng; g++) {
      const rsg_null_job &J = jobs[gfirst[(size_t)g]];
      rsg_feat_finish(&cnt[(size_t)g * 36], J.nleaf, J.alen, J.left, J.right, J.ld, J.rd, feats + (size_t)g * RSG_NULL_NFEAT);
    }
  }
  if (hists) {
    std::vector<HGeo> geo((size_t)ng);
    HIPCHK(hipMemcpy(geo.data(), d_geo, sizeof(HGeo) * ng, hipMemcpyDeviceToHost));
    std::vector<unsigned long long> h;
    for (int k = 0; k < ng; k++) {
      hists[k] = nullptr;
      const HGeo &g = geo[(size_t)k];
      if (!g.ok) continue;
      rsg_hist *H = rsg_hist_create(bmin, g.bmax, g.w);
      if (!H) return rsg_fail(errbuf, "out of memory");
      h.resize((size_t)g.nb);
      HIPCHK(hipMemcpy(h.data(), d_hist + g.off, sizeof(unsigned long long) * g.nb, hipMemcpyDeviceToHost));
      for (int b = 0; b < g.nb; b++) H->obs[b] = h[(size_t)b];
      H->n = H->Nc = H->No = (uint64_t)g.npairs;
      H->imin = H->cmin = g.imin; H->imax = g.imax; H->xmin = g.xmin; H->xmax = g.xmax;
      hists[k] = H;
    }
  }
  return RSG_OK;
}

extern "C" int rsg_ctx_graph_reuses(const rsg_ctx *c) { return c ? c->bexec_hits : 0; }

extern "C" int rsg_null_batch(rsg_ctx *c, const rsg_null_job *jobs, int nj, int covtype, double bmin, double pmass,
                              double fracfit, int doexpfit, rsg_hist **hists, double *fits, float *feats,
                              int use_graph, char *errbuf)
{
  if (!c || !jobs || !fits) return rsg_fail(errbuf, "rsg_null_batch: bad arguments");
  return null_batch_impl(c, jobs, nj, covtype, bmin, pmass, fracfit, doexpfit, hists, fits, feats, use_graph, nullptr,
                         nullptr, errbuf);
}

extern "C" int rsg_null_batch_fitch(rsg_ctx *c, const uint8_t *ax, int64_t nseq, int64_t alen, const int *left,
                                    const int *right, const uint64_t *seeds, int nrep, int covtype, double bmin,
                                    double pmass, double fracfit, int doexpfit, rsg_hist **hist, double *fit,
                                    int use_graph, char *errbuf)
{
  if (!c || !ax || !left || !right || !seeds || nrep < 1 || !fit) return rsg_fail(errbuf, "rsg_null_batch_fitch: bad arguments");
  FitchPlan P;
  if (rsg_fitch_plan(nseq, alen, left, right, P, errbuf) != RSG_OK) return RSG_FAIL;
  std::vector<rsg_null_job> jobs((size_t)nrep);
  for (int r = 0; r < nrep; r++) {
    rsg_null_job &J = jobs[(size_t)r];
    J.nleaf = nseq; J.alen = alen; J.left = left; J.right = right; J.ld = nullptr; J.rd = nullptr; J.model = nullptr;
    J.seed = seeds[r]; J.stream = (uint32_t)r; J.gapfrac = 0.; J.group = 0;
  }
  return null_batch_impl(c, jobs.data(), nrep, covtype, bmin, pmass, fracfit, doexpfit, hist, fit, nullptr, use_graph, ax,
                         &P, errbuf);
}
