// rsg_stats.hip -- probabilities, per-pair statistics, background corrections.
//
// Reference loops replaced (src/correlators.c, R-scape v2.6.11):
//   corr_Marginals :1337-1375, mutual_naive_psi :1782-1817, corr_ValidateProbs :1499-1545,
//   corr_Calculate{CHI,OMES,GT,MI,MIr,MIg}_{C16,C2,CWC} :49-874, CCF :984-1061,
//   RAF/RAFS :876-982, corr_CalculateCOVCorrected :1063-1157.
//
// Pair records hold raw counts c[x][y] (unit or FP32 weighted).  The reference's
// probabilities are pp = (c + 1e-10) / (nseff + 16e-10) with nseff = sum c, and
// obs = nseff pp, exp = nseff pm_i[x] pm_j[y]; every statistic is rewritten in terms of
// pp and q = pm_i pm_j (GT = 2 nseff sum pp ln(pp/q), CHI = nseff sum (pp-q)^2/q, ...),
// which is the same quantity with the same cell conditions.  In FP64 mode integer counts
// take ln(c + 1e-10) and ln(nseff + 16e-10) from per-alignment tables.
#include "rsg_device.h"
#include <cmath>
#include <algorithm>

static int ensure_red(rsg_ctx *c, size_t n, char *errbuf);
#define RSG_MINMAX_BLOCKS 512   // partials of k_minmax_upper (RAF / RAFS extremes)

// Reduction scratch large enough for every stage of one scoring of an L-column alignment
// (sized before a deferred chain, so no stage reallocates it while earlier work is queued).
int rsg_reserve_scratch(rsg_ctx *c, int64_t L, char *errbuf)
{
  const int64_t nbt = (L + 63) / 64, nb = (L + 31) / 32;
  const size_t need = std::max<size_t>({(size_t)(pm_npi(nbt) + pm_npj(nbt)), (size_t)(2 * pm32_n(nb)),
                                        (size_t)(2 * nb * (nb + 1) + 8 + 4 * L), (size_t)(2 * L + 8),
                                        (size_t)(2 * RSG_MINMAX_BLOCKS + 8)});
  return ensure_red(c, need, errbuf);
}

// ---- pm: corr_Marginals (pm_tile / pm_reduce_col in rsg_device.h).  Tiles below the
// diagonal exit at once.
template <int KIND>
__global__ void __launch_bounds__(64 * PMW) k_pm_tile(const void *__restrict__ cnt, int64_t L, int nbt,
                                                      double *__restrict__ part_i, double *__restrict__ part_j,
                                                      int *__restrict__ flag)
{
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *flag = 0;   // k_pm_reduce_ps sets it
  const int bj = (int)blockIdx.x / PMS, sub = (int)blockIdx.x % PMS, bi = (int)blockIdx.y;
  if (bi > bj) return;
  pm_tile<KIND>(cnt, L, nbt, bi, bj, sub, part_i, part_j);
}

// Scratch for the marginal partials the int8 Gram leaves (2 pm32_n doubles), in d_red.
double *rsg_pm32_scratch(rsg_ctx *c, int64_t L, char *errbuf)
{
  if (ensure_red(c, 2 * (size_t)pm32_n((L + 31) / 32), errbuf) != RSG_OK) return nullptr;
  return c->d_red;
}


// bytes of w equal to the byte pattern pat (0x01010101 x): a byte of w ^ pat is zero exactly
// there, and ((b & 0x7F) + 0x7F) | b has its high bit set exactly when b != 0
__device__ __forceinline__ unsigned bytes_eq(unsigned w, unsigned pat)
{
  const unsigned y = w ^ pat;
  return 4u - (unsigned)__popc((((y & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | y) & 0x80808080u);
}

// corr_NaivePS: ps[i][x] = normalize(1e-5 + sum_s w_s [a_si == x]), x over A, C, G, U, gap.
// Unit weights: one wave per column (4 per block), exact integer counts, 16 codes per load, four
// loads in flight per lane, the padding codes (4) of the last load taken off the gap count.
__device__ __forceinline__ void ps_cols(const uint8_t *__restrict__ axT, int64_t N, int64_t Npad, int64_t L,
                                        double *__restrict__ ps, int64_t blk)
{
  const int lane = threadIdx.x & 63;
  const int64_t i = blk * 4 + (threadIdx.x >> 6);
  if (i >= L) return;
  const uint4 *col = (const uint4 *)(axT + i * Npad);
  const int64_t nch = (N + 15) / 16;
  unsigned n[5] = {0u, 0u, 0u, 0u, 0u};
  auto count = [&](const uint4 v) {
#pragma unroll
    for (int x = 0; x < 5; x++)
      n[x] += bytes_eq(v.x, 0x01010101u * x) + bytes_eq(v.y, 0x01010101u * x) + bytes_eq(v.z, 0x01010101u * x) +
              bytes_eq(v.w, 0x01010101u * x);
  };
  int64_t k = lane;
  for (; k + 192 < nch; k += 256) {
    const uint4 v0 = col[k], v1 = col[k + 64], v2 = col[k + 128], v3 = col[k + 192];
    count(v0); count(v1); count(v2); count(v3);
  }
  for (; k < nch; k += 64) count(col[k]);
#pragma unroll
  for (int x = 0; x < 5; x++)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) n[x] += __shfl_xor(n[x], o, 64);
  if (lane == 0) {
    n[4] -= (unsigned)(16 * nch - N);
    double v[5], sum = 0.;
    for (int x = 0; x < 5; x++) { v[x] = 1e-5 + (double)n[x]; sum += v[x]; }
    for (int x = 0; x < 5; x++) ps[i * 5 + x] = v[x] / sum;
  }
}

// Weighted (the caller's work group covers one column, i; more groups in flight than with several
// columns per group, which measured slower): wave q of 4 takes sequences 4 l .. 4 l + 3 (l: lane) of every
// 1024 from 256 q on (one 32-bit word of codes and the four weights of consecutive sequences per
// step, so a wave's loads are contiguous), FP64 sums per lane, then the wave sums and the four
// waves' sums in a fixed order.
__device__ __forceinline__ void ps_col_w(const uint8_t *__restrict__ axT, const double *__restrict__ w, int64_t N,
                                         int64_t Npad, double *__restrict__ ps, int64_t i)
{
  __shared__ double sh[4][5];
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const uint32_t *c32 = (const uint32_t *)(axT + i * Npad);
  double acc[5] = {0., 0., 0., 0., 0.};
  for (int64_t s0 = 4 * (int64_t)lane + 256 * q; s0 < N; s0 += 1024) {
    const uint32_t c4 = c32[s0 >> 2];
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int64_t sq = s0 + b;
      const unsigned a = (c4 >> (8 * b)) & 0xFFu;
      const double ws = (sq < N && a < 5u) ? w[sq] : 0.0;
#pragma unroll
      for (int x = 0; x < 5; x++) acc[x] += (a == (unsigned)x) ? ws : 0.0;
    }
  }
#pragma unroll
  for (int x = 0; x < 5; x++)
    for (int o = 32; o > 0; o >>= 1) acc[x] += __shfl_down(acc[x], o, 64);
  if (lane == 0)
#pragma unroll
    for (int x = 0; x < 5; x++) sh[q][x] = acc[x];
  __syncthreads();
  if (threadIdx.x == 0) {
    double v[5], sum = 0.;
    for (int x = 0; x < 5; x++) { v[x] = 1e-5 + ((sh[0][x] + sh[1][x]) + (sh[2][x] + sh[3][x])); sum += v[x]; }
    for (int x = 0; x < 5; x++) ps[i * 5 + x] = v[x] / sum;
  }
}

// The marginals' second pass and the single-residue frequencies in one launch (they are
// independent): blocks [0, nbr) reduce four columns' marginal partials each (pm_reduce_col),
// the rest compute ps, one column per block weighted (ps_col_w), four per block unweighted
// (ps_cols).
template <bool W>
__global__ void __launch_bounds__(256) k_pm_reduce_ps(const double *__restrict__ part_i, const double *__restrict__ part_j,
                                                      int nbt, int64_t L, double tol, double *__restrict__ pm,
                                                      double *__restrict__ lpm, int *__restrict__ flag, int nbr,
                                                      const uint8_t *__restrict__ axT, const double *__restrict__ w,
                                                      int64_t N, int64_t Npad, double *__restrict__ ps)
{
  if ((int)blockIdx.x < nbr) {
    const int64_t col = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (col < L) pm_reduce_col(part_i, part_j, nbt, L, col, threadIdx.x & 63, tol, pm, lpm, flag);
    return;
  }
  const int64_t b = (int64_t)blockIdx.x - nbr;
  if (W) ps_col_w(axT, w, N, Npad, ps, b);
  else   ps_cols(axT, N, Npad, L, ps, b);
}

// Unit weights after the int8 Gram, in one launch: the marginals from the Gram's 32 x 32 block
// partials (blocks [0, nbm), 16 lanes per column) and the single-residue frequencies (the rest,
// ps_cols).
__global__ void __launch_bounds__(256) k_pm32_ps(const double *__restrict__ part, int nb32, int64_t L, double tol,
                                                 double *__restrict__ pm, double *__restrict__ lpm, int *__restrict__ flag,
                                                 const uint8_t *__restrict__ axT, int64_t N, int64_t Npad,
                                                 double *__restrict__ ps, int nbm)
{
  if ((int)blockIdx.x < nbm) {
    const int64_t col = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    if (col < L) pm32_reduce_col(part, part + pm32_n(nb32), nb32, L, col, threadIdx.x & 15, tol, pm, lpm, flag);
    return;
  }
  ps_cols(axT, N, Npad, L, ps, (int64_t)blockIdx.x - nbm);
}

// ln(k + 1e-10), ln(k + 16e-10) and (k + 1e-10) ln(k + 1e-10) for integer k in [0, N]
__global__ void k_logtab(int64_t N, double *__restrict__ t)
{
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k <= N) {
    const double l = log((double)k + 1e-10);
    t[k] = l; t[N + 1 + k] = log((double)k + 16e-10); t[2 * (N + 1) + k] = ((double)k + 1e-10) * l;
  }
}

int rsg_probs_from_counts(rsg_ctx *c, rsg_msa *m, double tol, char *errbuf)
{
  const int64_t L = m->L;
  if (!m->d_pm) HIPCHK(hipMalloc(&m->d_pm, sizeof(double) * 8 * (size_t)L));
  if (!m->d_ps) HIPCHK(hipMalloc(&m->d_ps, sizeof(double) * 5 * (size_t)L));
  int t;
  if (m->pm_fused) {                 // unit weights (int8 counts)
    // the Gram left 32 x 32 block partials in d_red (rsg_pm32_scratch) and cleared the flag
    m->pm_fused = 0;
    const int nb32 = (int)((L + 31) / 32);
    const int nbm = (int)((L + 15) / 16), nbp = (int)((L + 3) / 4);
    rsg_tstart(c, "marginals+ps", &t);
    hipLaunchKernelGGL(k_pm32_ps, dim3((unsigned)(nbm + nbp)), dim3(256), 0, c->stream, c->d_red, nb32, L, tol,
                       m->d_pm, m->d_pm + 4 * L, (int *)(c->d_res + 4), m->d_axT, m->N, m->Npad, m->d_ps, nbm);
    rsg_tstop(c, t);
    HIPCHK(hipGetLastError());
  } else {
  const int nbt = (int)((L + PMT - 1) / PMT);
  if (ensure_red(c, (size_t)(pm_npi(nbt) + pm_npj(nbt)), errbuf) != RSG_OK) return RSG_FAIL;
  double *part_i = c->d_red, *part_j = c->d_red + pm_npi(nbt);
  rsg_tstart(c, "marginals+ps", &t);
  if (m->cnt_kind == CNT_U16)
    hipLaunchKernelGGL(k_pm_tile<CNT_U16>, dim3((unsigned)(nbt * PMS), (unsigned)nbt), dim3(64 * PMW), 0, c->stream, m->d_cnt, L, nbt, part_i, part_j,
                       (int *)(c->d_res + 4));
  else if (m->cnt_kind == CNT_U32)
    hipLaunchKernelGGL(k_pm_tile<CNT_U32>, dim3((unsigned)(nbt * PMS), (unsigned)nbt), dim3(64 * PMW), 0, c->stream, m->d_cnt, L, nbt, part_i, part_j,
                       (int *)(c->d_res + 4));
  else
    hipLaunchKernelGGL(k_pm_tile<CNT_F32>, dim3((unsigned)(nbt * PMS), (unsigned)nbt), dim3(64 * PMW), 0, c->stream, m->d_cnt, L, nbt, part_i, part_j,
                       (int *)(c->d_res + 4));
  // the second pass of the marginals with the single-residue frequencies, in one launch
  const int nbr = (int)((L + 3) / 4), nps = m->weighted ? (int)L : (int)((L + 3) / 4);
  if (m->weighted)
    hipLaunchKernelGGL(k_pm_reduce_ps<true>, dim3((unsigned)(nbr + nps)), dim3(256), 0, c->stream, part_i, part_j, nbt, L, tol,
                       m->d_pm, m->d_pm + 4 * L, (int *)(c->d_res + 4), nbr, m->d_axT, m->d_w, m->N, m->Npad, m->d_ps);
  else
    hipLaunchKernelGGL(k_pm_reduce_ps<false>, dim3((unsigned)(nbr + nps)), dim3(256), 0, c->stream, part_i, part_j, nbt, L, tol,
                       m->d_pm, m->d_pm + 4 * L, (int *)(c->d_res + 4), nbr, m->d_axT, m->d_w, m->N, m->Npad, m->d_ps);
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  }
  if (!m->d_lgc) HIPCHK(hipMalloc(&m->d_lgc, sizeof(double) * 3 * (size_t)(m->N + 1)));
  if (m->lgc_n != m->N) {
    hipLaunchKernelGGL(k_logtab, dim3((unsigned)((m->N + 256) / 256)), dim3(256), 0, c->stream, m->N, m->d_lgc);
    HIPCHK(hipGetLastError());
    m->lgc_n = m->N;
  }
  return RSG_OK;
}

// ------------------------------------------------------------------------------------
// One 32 x 32 pair tile on or above the diagonal per block (nb (nb + 1) / 2 blocks); COV
// written to both triangles through an LDS transpose.
template <typename T, int KIND, bool TAB, int TYPE = -1, int CLS = -1>
__global__ void __launch_bounds__(256) k_stat(StatArgs A, int nb)
{
  int bi, bj;
  tri_tile(blockIdx.x, bi, bj);
  stat_tile<T, KIND, TAB, TYPE, CLS>(A, bi, bj, (int64_t)blockIdx.x);
}

// FP64 statistics from integer counts with the statistic and class fixed at compile time
// (the common cases: fewer live registers, no dead branches).  The global log table beats
// inline FP64 log() by 2x (45 vs 99 us for GT at L = 1500) and a per-block LDS copy of it
// (49 us).  Returns the number of min / max partials written (0: not handled here).
template <int TYPE>
static int launch_stat_fixed(rsg_ctx *c, const StatArgs &A, int cls, int nb)
{
  if (cls != RSG_C16 && cls != RSG_C2) return 0;
  const int ntri = nb * (nb + 1) / 2;
  const dim3 grid((unsigned)ntri);
  if (cls == RSG_C16) hipLaunchKernelGGL((k_stat<double, CNT_U16, true, TYPE, RSG_C16>), grid, dim3(256), 0, c->stream, A, nb);
  else                hipLaunchKernelGGL((k_stat<double, CNT_U16, true, TYPE, RSG_C2>), grid, dim3(256), 0, c->stream, A, nb);
  return ntri;
}

// FP64 statistics of class C16 from FP32 (weighted) records with the statistic fixed at compile
// time (inline log: non-integer counts have no table); 0 when not handled here.
template <int TYPE>
static int launch_stat_f64w_fixed(rsg_ctx *c, const StatArgs &A, int cls, int nb)
{
  if (cls != RSG_C16) return 0;
  const int ntri = nb * (nb + 1) / 2;
  hipLaunchKernelGGL((k_stat<double, CNT_F32, false, TYPE, RSG_C16>), dim3((unsigned)ntri), dim3(256), 0, c->stream, A, nb);
  return ntri;
}

// FP32 short-form statistics over half tiles: blocks 2t and 2t + 1 take the two halves of pair
// tile t (bi <= bj); block min / max to part.
template <int KIND, int TYPE>
__global__ void __launch_bounds__(256) k_stat_half(StatArgs A, int nb)
{
  int bi, bj;
  tri_tile(blockIdx.x >> 1, bi, bj);
  double lmin = INFINITY, lmax = -INFINITY;
  stat_half_f32c16<KIND, TYPE>(A, bi, bj, (int)(blockIdx.x & 1), lmin, lmax);
  block_minmax(lmin, lmax, A.part, (int64_t)blockIdx.x);
}

// FP32 statistics of class C16 with the statistic fixed at compile time, from uint16 or FP32
// (weighted) records, in the short form over half tiles (pair_stat_f32c16); 0 when not handled
// here (other classes go through pair_stat_f32).
template <int TYPE>
static int launch_stat_f32_fixed(rsg_ctx *c, const StatArgs &A, int kind, int cls, int nb)
{
  if (cls != RSG_C16 || (kind != CNT_U16 && kind != CNT_F32)) return 0;
  const int ntri = nb * (nb + 1) / 2;
  if constexpr (TYPE == RSG_GT || TYPE == RSG_MI || TYPE == RSG_MIg || TYPE == RSG_MIr || TYPE == RSG_CHI || TYPE == RSG_OMES) {
    const dim3 grid((unsigned)(2 * ntri));
    if (kind == CNT_U16) hipLaunchKernelGGL((k_stat_half<CNT_U16, TYPE>), grid, dim3(256), 0, c->stream, A, nb);
    else                 hipLaunchKernelGGL((k_stat_half<CNT_F32, TYPE>), grid, dim3(256), 0, c->stream, A, nb);
    return 2 * ntri;
  } else {
    const dim3 grid((unsigned)ntri);
    if (kind == CNT_U16) hipLaunchKernelGGL((k_stat<float, CNT_U16, false, TYPE, RSG_C16>), grid, dim3(256), 0, c->stream, A, nb);
    else                 hipLaunchKernelGGL((k_stat<float, CNT_F32, false, TYPE, RSG_C16>), grid, dim3(256), 0, c->stream, A, nb);
    return ntri;
  }
}

// (min, max) of n block partials into out[0..1]; one block of 256 threads, four partials per
// thread loaded at once, then block_minmax256.  With hout != nullptr out[0..4] (min, max and the
// flags and COVavg the chain left in out[2..4]) also go to hout, host memory, followed (after a
// system-scope release) by the completion stamp in hout[63], which the caller waits for.
__global__ void __launch_bounds__(256) k_reduce_minmax(const double *__restrict__ part, int n, double *__restrict__ out,
                                                       double *__restrict__ hout, unsigned long long stamp)
{
  __shared__ double rmin[256], rmax[256];
  double a = INFINITY, b = -INFINITY;
  for (int k0 = threadIdx.x; k0 < n; k0 += 1024) {
    double2 v[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int k = k0 + 256 * q;
      v[q] = (k < n) ? *(const double2 *)(part + 2 * k) : make_double2(INFINITY, -INFINITY);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) { a = fmin(a, v[q].x); b = fmax(b, v[q].y); }
  }
  block_minmax256(a, b, rmin, rmax);
  if (threadIdx.x == 0) {
    out[0] = a; out[1] = b;
    if (hout) {
      hout[0] = a; hout[1] = b;
      hout[2] = out[2]; hout[3] = out[3]; hout[4] = out[4];
      __threadfence_system();
      *(volatile unsigned long long *)(hout + 63) = stamp;
    }
  }
}

static int ensure_red(rsg_ctx *c, size_t n, char *errbuf)
{
  if (c->red_cap >= n) return RSG_OK;
  if (c->no_alloc) return rsg_fail(errbuf, "reduction scratch of %zu doubles exceeds the fixed %zu", n, c->red_cap);
  if (c->d_red) hipFree(c->d_red);
  c->d_red = nullptr;
  HIPCHK(hipMalloc(&c->d_red, sizeof(double) * n));
  c->red_cap = n;
  return RSG_OK;
}

// Block partials -> (min, max) into d_res[0..1] (and h_pinned[0..1] unless the chain is
// deferred: rsg_cov_Calculate copies d_res once at its end).
static int fetch_minmax(rsg_ctx *c, rsg_msa *m, int nparts, char *errbuf)
{
  if (c->skip_minmax) return RSG_OK;            // a correction recomputes them
  // a deferred chain ends here: the reduction also writes the result block to host memory
  hipLaunchKernelGGL(k_reduce_minmax, dim3(1), dim3(256), 0, c->stream, c->d_red, nparts, c->d_res,
                     c->defer ? c->d_hres : nullptr, ++c->hstamp);
  HIPCHK(hipGetLastError());
  if (c->defer) { c->res_on_host = 1; return RSG_OK; }
  HIPCHK(hipMemcpyAsync(c->h_pinned, c->d_res, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  m->minCOV = c->h_pinned[0]; m->maxCOV = c->h_pinned[1];
  return RSG_OK;
}

// CCF global meanp (src/correlators.c:1027-1036): sum over i<j of nseff_ij pm_i[x], normalised.
template <int KIND>
__global__ void k_ccf_rows(const void *__restrict__ cnt, int64_t L, const double *__restrict__ pm, double *__restrict__ rows)
{
  __shared__ double red[4][256];
  const int64_t i = blockIdx.x;
  double ns_sum = 0.;
  for (int64_t j = i + 1 + threadIdx.x; j < L; j += 256) {
    double cd[16]; load_rec<KIND>(cnt, rsg_rec(i, j, L), cd);
    for (int k = 0; k < 16; k++) ns_sum += cd[k];
  }
  for (int x = 0; x < 4; x++) red[x][threadIdx.x] = ns_sum * pm[i * 4 + x];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) for (int x = 0; x < 4; x++) red[x][threadIdx.x] += red[x][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) for (int x = 0; x < 4; x++) rows[i * 4 + x] = red[x][0];
}

int rsg_stat(rsg_ctx *c, rsg_msa *m, int base, int cls, const double *allow16, int prec, char *errbuf)
{
  const int64_t L = m->L;
  if (!m->d_cov) HIPCHK(hipMalloc(&m->d_cov, sizeof(double) * (size_t)(L * L)));
  const int nb = (int)((L + 31) / 32);
  const int nblk = nb * (nb + 1) / 2;             // pair tiles on or above the diagonal
  // block min / max partials (2 per block; half-tile kernels have 2 nblk blocks), CCF rows after
  if (ensure_red(c, 4 * (size_t)nblk + 8 + 4 * (size_t)L, errbuf) != RSG_OK) return RSG_FAIL;
  StatArgs A;
  A.cnt = m->d_cnt; A.pm = m->d_pm; A.lpm = m->d_pm + 4 * L; A.lgc = m->d_lgc;
  A.L = L; A.N = m->N; A.wsum = rsg_count_wsum(m); A.cov = m->d_cov; A.part = c->d_red;
  A.type = base; A.cls = cls;
  A.nsx = m->wq_counts ? m->d_nsx : nullptr;
  A.allow = 0;
  for (int k = 0; k < 16; k++) if (allow16[k] > 0) A.allow |= 1u << k;
  for (int x = 0; x < 4; x++) A.meanp[x] = 0.;
  if (base == RSG_CCF) {
    double *rows = c->d_red + 2 * (size_t)nblk + 8;
    if (m->cnt_kind == CNT_U16) hipLaunchKernelGGL(k_ccf_rows<CNT_U16>, dim3((unsigned)L), dim3(256), 0, c->stream, m->d_cnt, L, m->d_pm, rows);
    else if (m->cnt_kind == CNT_U32) hipLaunchKernelGGL(k_ccf_rows<CNT_U32>, dim3((unsigned)L), dim3(256), 0, c->stream, m->d_cnt, L, m->d_pm, rows);
    else                            hipLaunchKernelGGL(k_ccf_rows<CNT_F32>, dim3((unsigned)L), dim3(256), 0, c->stream, m->d_cnt, L, m->d_pm, rows);
    std::vector<double> h(4 * (size_t)L);
    HIPCHK(hipMemcpyAsync(h.data(), rows, sizeof(double) * 4 * (size_t)L, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double mp[4] = {0., 0., 0., 0.}, s = 0.;
    for (int64_t i = 0; i < L; i++) for (int x = 0; x < 4; x++) mp[x] += h[i * 4 + x];
    for (int x = 0; x < 4; x++) s += mp[x];
    for (int x = 0; x < 4; x++) A.meanp[x] = (s != 0.) ? mp[x] / s : 0.25;
  }
  const bool tab = (m->cnt_kind != CNT_F32) && prec == RSG_STAT_FP64;
  int t; rsg_tstart(c, "stat", &t);
  dim3 grid((unsigned)nblk), blk(256);
#define LAUNCH(T, K, TB) hipLaunchKernelGGL((k_stat<T, K, TB>), grid, blk, 0, c->stream, A, nb)
  int nparts = 0;
  if (prec == RSG_STAT_FP64 && m->cnt_kind == CNT_U16 && tab) {
    switch (base) {
    case RSG_GT:   nparts = launch_stat_fixed<RSG_GT>(c, A, cls, nb); break;
    case RSG_MI:   nparts = launch_stat_fixed<RSG_MI>(c, A, cls, nb); break;
    case RSG_MIr:  nparts = launch_stat_fixed<RSG_MIr>(c, A, cls, nb); break;
    case RSG_MIg:  nparts = launch_stat_fixed<RSG_MIg>(c, A, cls, nb); break;
    case RSG_CHI:  nparts = launch_stat_fixed<RSG_CHI>(c, A, cls, nb); break;
    case RSG_OMES: nparts = launch_stat_fixed<RSG_OMES>(c, A, cls, nb); break;
    }
  }
  if (prec == RSG_STAT_FP64 && m->cnt_kind == CNT_F32) {
    switch (base) {
    case RSG_GT:   nparts = launch_stat_f64w_fixed<RSG_GT>(c, A, cls, nb); break;
    case RSG_MI:   nparts = launch_stat_f64w_fixed<RSG_MI>(c, A, cls, nb); break;
    case RSG_MIr:  nparts = launch_stat_f64w_fixed<RSG_MIr>(c, A, cls, nb); break;
    case RSG_MIg:  nparts = launch_stat_f64w_fixed<RSG_MIg>(c, A, cls, nb); break;
    case RSG_CHI:  nparts = launch_stat_f64w_fixed<RSG_CHI>(c, A, cls, nb); break;
    }
  }
  if (prec == RSG_STAT_FP32) {
    switch (base) {
    case RSG_GT:   nparts = launch_stat_f32_fixed<RSG_GT>(c, A, m->cnt_kind, cls, nb); break;
    case RSG_MI:   nparts = launch_stat_f32_fixed<RSG_MI>(c, A, m->cnt_kind, cls, nb); break;
    case RSG_MIr:  nparts = launch_stat_f32_fixed<RSG_MIr>(c, A, m->cnt_kind, cls, nb); break;
    case RSG_MIg:  nparts = launch_stat_f32_fixed<RSG_MIg>(c, A, m->cnt_kind, cls, nb); break;
    case RSG_CHI:  nparts = launch_stat_f32_fixed<RSG_CHI>(c, A, m->cnt_kind, cls, nb); break;
    case RSG_OMES: nparts = launch_stat_f32_fixed<RSG_OMES>(c, A, m->cnt_kind, cls, nb); break;
    }
  }
  if (nparts > 0) {
  } else if (prec == RSG_STAT_FP64) {
    if (m->cnt_kind == CNT_U16) { if (tab) LAUNCH(double, CNT_U16, true); else LAUNCH(double, CNT_U16, false); }
    else if (m->cnt_kind == CNT_U32) LAUNCH(double, CNT_U32, true);
    else                        LAUNCH(double, CNT_F32, false);
  } else {
    if (m->cnt_kind == CNT_U16) LAUNCH(float, CNT_U16, false);
    else if (m->cnt_kind == CNT_U32) LAUNCH(float, CNT_U32, false);
    else                        LAUNCH(float, CNT_F32, false);
  }
#undef LAUNCH
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  if (fetch_minmax(c, m, nparts > 0 ? nparts : nblk, errbuf) != RSG_OK) return RSG_FAIL;
  m->have_cov = 1;
  return RSG_OK;
}

// ------------------------------------------------------------------------------------
// RAF from unit-weight counts (exact integer algebra; see rsg_capi for the derivation),
// RAFS as the 3-point anti-diagonal stencil (src/correlators.c:954-960).
template <int KIND>
__global__ void __launch_bounds__(256) k_raf(const void *__restrict__ cnt, int64_t L, int64_t N, unsigned allow,
                                             double *__restrict__ cov)
{
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= L * L) return;
  int64_t i = idx / L, j = idx - i * L;
  if (i == j) { cov[idx] = -INFINITY; return; }
  int64_t a = (i < j) ? i : j, b = (i < j) ? j : i;
  double cd[16]; load_rec<KIND>(cnt, rsg_rec(a, b, L), cd);
  double M = 0., mx[4] = {0, 0, 0, 0}, my[4] = {0, 0, 0, 0};
  for (int x = 0; x < 4; x++)
    for (int y = 0; y < 4; y++)
      if ((allow >> (x * 4 + y)) & 1u) { double n = cd[x * 4 + y]; M += n; mx[x] += n; my[y] += n; }
  double sx = 0., sy = 0.;
  for (int k = 0; k < 4; k++) { sx += mx[k] * mx[k]; sy += my[k] * my[k]; }
  double cij = 0.5 * (M * M - sx) + 0.5 * (M * M - sy);
  cij /= (N > 1) ? (double)N * ((double)N - 1.0) : 1.0;
  cij *= 2.0;
  double qij = ((double)N - M) / (double)N;
  cov[idx] = cij - qij;
}

__global__ void k_rafs(const double *__restrict__ B, int64_t L, double *__restrict__ cov)
{
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= L * L) return;
  int64_t r = idx / L, s = idx - r * L;
  if (r == s) { cov[idx] = -INFINITY; return; }
  int64_t i = (r < s) ? r : s, j = (r < s) ? s : r;
  double v = 2.0 * B[i * L + j];
  if (i > 0 && j < L - 1)              v += B[(i - 1) * L + (j + 1)];
  if (j > 0 && i < L - 1 && i < j - 2) v += B[(i + 1) * L + (j - 1)];
  cov[idx] = 0.25 * v;
}

template <int KIND>
__global__ void k_minmax_upper(const double *__restrict__ cov, int64_t L, double *__restrict__ part)
{
  __shared__ double rmin[256], rmax[256];
  double a = INFINITY, b = -INFINITY;
  int64_t n = L * L;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * 256) {
    int64_t i = idx / L, j = idx - i * L;
    if (i < j) { double v = cov[idx]; a = fmin(a, v); b = fmax(b, v); }
  }
  rmin[threadIdx.x] = a; rmax[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) { rmin[threadIdx.x] = fmin(rmin[threadIdx.x], rmin[threadIdx.x + s]); rmax[threadIdx.x] = fmax(rmax[threadIdx.x], rmax[threadIdx.x + s]); }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = rmin[0]; part[2 * blockIdx.x + 1] = rmax[0]; }
}

int rsg_minmax(rsg_ctx *c, rsg_msa *m, char *errbuf)
{
  const int nblk = RSG_MINMAX_BLOCKS;
  if (ensure_red(c, 2 * nblk + 8, errbuf) != RSG_OK) return RSG_FAIL;
  hipLaunchKernelGGL(k_minmax_upper<0>, dim3(nblk), dim3(256), 0, c->stream, m->d_cov, m->L, c->d_red);
  HIPCHK(hipGetLastError());
  return fetch_minmax(c, m, nblk, errbuf);
}

int rsg_raf(rsg_ctx *c, rsg_msa *m, int rafs, const double *allow16, char *errbuf)
{
  const int64_t L = m->L;
  if (!m->d_cov) HIPCHK(hipMalloc(&m->d_cov, sizeof(double) * (size_t)(L * L)));
  unsigned allow = 0;
  for (int k = 0; k < 16; k++) if (allow16[k] > 0) allow |= 1u << k;
  const void *cnt = m->d_cnt_u ? m->d_cnt_u : m->d_cnt;
  int kind = m->d_cnt_u ? m->cnt_u_kind : m->cnt_kind;
  if (kind == CNT_F32) return rsg_fail(errbuf, "RAF needs unit-weight counts");
  double *dst = m->d_cov;
  if (rafs) {
    if (!m->d_tmp) HIPCHK(hipMalloc(&m->d_tmp, sizeof(double) * (size_t)(L * L)));
    dst = m->d_tmp;
  }
  unsigned nb = (unsigned)((L * L + 255) / 256);
  int t; rsg_tstart(c, "raf", &t);
  if (kind == CNT_U16) hipLaunchKernelGGL(k_raf<CNT_U16>, dim3(nb), dim3(256), 0, c->stream, cnt, L, m->N, allow, dst);
  else                 hipLaunchKernelGGL(k_raf<CNT_U32>, dim3(nb), dim3(256), 0, c->stream, cnt, L, m->N, allow, dst);
  if (rafs) hipLaunchKernelGGL(k_rafs, dim3(nb), dim3(256), 0, c->stream, m->d_tmp, L, m->d_cov);
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  if (rsg_minmax(c, m, errbuf) != RSG_OK) return RSG_FAIL;
  m->have_cov = 1;
  return RSG_OK;
}

// ------------------------------------------------------------------------------------
// corr_CalculateCOVCorrected (src/covariation.c): row sums over j != i (FP64, fixed order) and
// the row means r / (L - 1), COVavg, then the APC / ASC correction in place with a NaN flag and
// block min / max partials (k_reduce_minmax finishes them).
__device__ __forceinline__ double row_dl(int64_t L) { return (L > 1) ? (double)L - 1. : 1.; }

// Row sums from COV, one block per row (the order k_rowsum_b of the batched path reproduces).
__global__ void __launch_bounds__(256) k_rowsum(const double *__restrict__ cov, int64_t L, double *__restrict__ rows,
                                                int *__restrict__ flag)
{
  __shared__ double red[257];
  const int64_t i = blockIdx.x;
  if (i == 0 && threadIdx.x == 0) *flag = 0;       // NaN flag of k_correct (runs after this kernel)
  double s = 0.;
  for (int64_t j = threadIdx.x; j < L; j += 256) if (j != i) s += cov[i * L + j];
  s = block_sum_tree256(s, red);
  if (threadIdx.x == 0) { rows[L + i] = s; rows[i] = s / row_dl(L); }
}

// APC / ASC in place, one block per row i (no per-element index division), four elements per
// thread in flight; block min / max of the corrected row to part[2 i], part[2 i + 1].  Every
// block forms COVavg from the raw row sums rows[L..2L) with the loop and tree of the batched
// k_rowmean_b (same order, same value; block 0 publishes it); the row means come from rows[0..L).
__global__ void __launch_bounds__(256) k_correct(double *__restrict__ cov, int64_t L, const double *__restrict__ rows,
                                                 int actype, int *__restrict__ flag, double *__restrict__ part,
                                                 double *__restrict__ avg_out)
{
  __shared__ double rmin[257], rmax[256];
  const int64_t i = blockIdx.x;
  double *C = cov + i * L;
  // the row's first four elements per thread are loaded before COVavg is formed, so their
  // latency overlaps the reduction
  double v[4], xj[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int64_t j = threadIdx.x + 256 * q;
    if (j < L) { v[q] = C[j]; xj[q] = rows[j]; }
  }
  double s = 0.;
  for (int64_t k = threadIdx.x; k < L; k += 256) s += rows[L + k];
  s = block_sum_tree256(s, rmin);
  const double avg = (L > 1) ? s / ((double)L * ((double)L - 1.)) : s;
  __syncthreads();                                   // rmin is reused below
  if (blockIdx.x == 0 && threadIdx.x == 0) *avg_out = avg;
  const double xi = rows[i];
  double a = INFINITY, b = -INFINITY;
  for (int64_t j0 = threadIdx.x; j0 < L; j0 += 1024) {
    if (j0 != threadIdx.x) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int64_t j = j0 + 256 * q;
        if (j < L) { v[q] = C[j]; xj[q] = rows[j]; }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int64_t j = j0 + 256 * q;
      if (j < L && j != i) {
        const double r = (actype == RSG_APC) ? ((avg != 0.0) ? v[q] - xi * xj[q] / avg : 0.0) : v[q] - (xi + xj[q] - avg);
        if (isnan(r)) atomicOr(flag, 1);
        C[j] = r;
        a = fmin(a, r); b = fmax(b, r);
      }
    }
  }
  block_minmax256(a, b, rmin, rmax);
  if (threadIdx.x == 0) { part[2 * i] = a; part[2 * i + 1] = b; }
}

__global__ void k_shift(double *__restrict__ cov, int64_t L, const double *__restrict__ mm)
{
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= L * L) return;
  int64_t i = idx / L, j = idx - i * L;
  if (i != j) cov[idx] -= mm[0];
}

int rsg_correct(rsg_ctx *c, rsg_msa *m, int actype, int shiftnonneg, char *errbuf)
{
  const int64_t L = m->L;
  if (!m->have_cov) return rsg_fail(errbuf, "no covariation matrix to correct");
  if (!m->d_rowx) HIPCHK(hipMalloc(&m->d_rowx, sizeof(double) * (2 * (size_t)L + 2)));
  const int nblk = (int)L;
  if (ensure_red(c, 2 * (size_t)nblk + 8, errbuf) != RSG_OK) return RSG_FAIL;
  int *nanflag = (int *)(c->d_res + 2);
  // deferred chain without a shift: the min / max reduction also writes the result block to host
  // memory, so no copy follows
  const bool direct = c->defer && !shiftnonneg;
  int t; rsg_tstart(c, "correct", &t);
  hipLaunchKernelGGL(k_rowsum, dim3((unsigned)L), dim3(256), 0, c->stream, m->d_cov, L, m->d_rowx, nanflag);
  hipLaunchKernelGGL(k_correct, dim3(nblk), dim3(256), 0, c->stream, m->d_cov, L, m->d_rowx, actype, nanflag, c->d_red,
                     c->d_res + 3);
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_minmax, dim3(1), dim3(256), 0, c->stream, c->d_red, nblk, c->d_res, direct ? c->d_hres : nullptr,
                     ++c->hstamp);
  if (direct) { c->res_on_host = 1; return RSG_OK; }
  if (shiftnonneg)
    hipLaunchKernelGGL(k_shift, dim3((unsigned)((L * L + 255) / 256)), dim3(256), 0, c->stream, m->d_cov, L, c->d_res);
  HIPCHK(hipGetLastError());
  if (c->defer) return RSG_OK;
  HIPCHK(hipMemcpyAsync(c->h_pinned, c->d_res, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  int flag = *(int *)(c->h_pinned + 2);
  if (flag) return rsg_fail(errbuf, "bad covariation COVavg %f", c->h_pinned[3]);
  m->minCOV = c->h_pinned[0]; m->maxCOV = c->h_pinned[1];
  return RSG_OK;
}
