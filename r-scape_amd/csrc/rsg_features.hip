// rsg_features.hip -- the RSG_NULL_NFEAT (64) alignment + tree features the null-tail surrogate reads.
// No reference counterpart.
//
//   f[0]  ln N            f[1]  ln L            f[2]  ln(total branch length)
//   f[3..18]   column-entropy histogram: fraction of columns whose canonical-residue
//              entropy H (bits, 0..2) falls in bin min(15, floor(8 H))
//   f[19..34]  gap histogram: fraction of columns whose non-canonical fraction g falls in
//              bin min(15, floor(16 g))
//   f[35..37]  residue composition A, C, G (U is the rest)
//   f[38..47]  tree shape: mean branch length; its coefficient of variation; mean
//              root-to-leaf depth; depth CV; max depth / mean depth; mean terminal branch /
//              mean branch; mean internal branch / mean branch; mean root-to-leaf edge count /
//              log2 N; cherries / (N/2); Colless imbalance / ((N-1)(N-2)/2)
//   f[48..63]  split spectrum: every branch splits the leaves into its clade (n leaves) and the
//              rest; with the minor side's share q = min(n, N - n) / N in bin k = 0..7
//              (q in (1/4, 1/2], (1/8, 1/4], ..., (1/256, 1/128], then <= 1/256),
//              f[48 + k] = ln(1e-3 + sum of those branches' lengths t) and
//              f[56 + k] = ln(1e-4 + sum of (1 - e^-t)^2).  Two independent columns look
//              covarying when both change on one branch (probability ~ (1 - e^-t)^2), and how
//              far the score goes grows with the split's minor share; the tail of the null
//              follows the long branches above large clades, which the moments above miss.
//
// The column statistics are one pass over the one-hot Gram operand on the device (one wave
// per column; block histograms in LDS, one atomic per bin per block); the tree moments are
// O(N) on the host.
#include "rsg_internal.h"
#include <cmath>
#include <cstring>
#include <vector>

// Per item descriptor {one-hot panel offset, N, Npad, L, first block, Lpad}; the panel is
// the Gram operand (K-chunked [Npad / 16][Lpad][16] bytes, 1 << symbol for canonical
// residues, 0 otherwise).  One wave per column (4 per block) reads the column's Npad / 16
// chunks; 36 uint64 counters per item: [0..15] entropy bins, [16..31] gap bins, [32..35]
// residue totals.
__global__ void __launch_bounds__(256)
k_colfeat(const int64_t *__restrict__ items, int nitems, const uint8_t *__restrict__ base,
          unsigned long long *__restrict__ out)
{
  __shared__ unsigned int eh[16], gh[16];
  __shared__ unsigned long long comp[4];
  int lo = 0, hi = nitems - 1;
  while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (items[mid * 6 + 4] <= (int64_t)blockIdx.x) lo = mid; else hi = mid - 1; }
  const int64_t *it = items + lo * 6;
  const int64_t N = it[1], Npad = it[2], L = it[3], Lpad = it[5];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t c = ((int64_t)blockIdx.x - it[4]) * 4 + wave;
  if (threadIdx.x < 16) { eh[threadIdx.x] = 0; gh[threadIdx.x] = 0; }
  if (threadIdx.x < 4) comp[threadIdx.x] = 0;
  __syncthreads();
  if (c < L) {
    const uint4 *col = (const uint4 *)(base + it[0]) + c;
    unsigned int n[4] = {0, 0, 0, 0};
    for (int64_t k = lane; k < Npad / 16; k += 64) {
      const uint4 v = col[k * Lpad];
#pragma unroll
      for (int x = 0; x < 4; x++) {
        const unsigned mk = 0x01010101u << x;
        n[x] += __popc(v.x & mk) + __popc(v.y & mk) + __popc(v.z & mk) + __popc(v.w & mk);
      }
    }
#pragma unroll
    for (int x = 0; x < 4; x++)
      for (int o = 32; o > 0; o >>= 1) n[x] += __shfl_down(n[x], o, 64);
    if (lane == 0) {
      const unsigned int nc = n[0] + n[1] + n[2] + n[3];
      double H = 0.;
      for (int x = 0; x < 4; x++)
        if (n[x]) { const double p = (double)n[x] / (double)nc; H -= p * log2(p); }
      const double g = (double)(N - nc) / (double)N;
      atomicAdd(&eh[min(15, (int)(H * 8.0))], 1u);
      atomicAdd(&gh[min(15, (int)(g * 16.0))], 1u);
      for (int x = 0; x < 4; x++) atomicAdd(&comp[x], (unsigned long long)n[x]);
    }
  }
  __syncthreads();
  unsigned long long *o = out + (int64_t)lo * 36;
  if (threadIdx.x < 16) {
    if (eh[threadIdx.x]) atomicAdd(&o[threadIdx.x], (unsigned long long)eh[threadIdx.x]);
    if (gh[threadIdx.x]) atomicAdd(&o[16 + threadIdx.x], (unsigned long long)gh[threadIdx.x]);
  }
  if (threadIdx.x < 4 && comp[threadIdx.x]) atomicAdd(&o[32 + threadIdx.x], comp[threadIdx.x]);
}

__global__ void k_fill_u64f(unsigned long long *__restrict__ p, int64_t n)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0ull;
}

int rsg_feat_launch(rsg_ctx *c, const int64_t *d_items, int nitems, int64_t nblocks, const uint8_t *d_base,
                    unsigned long long *d_out)
{
  hipLaunchKernelGGL(k_fill_u64f, dim3((unsigned)((nitems * 36 + 255) / 256)), dim3(256), 0, c->stream, d_out,
                     (int64_t)nitems * 36);
  if (nblocks > 0)
    hipLaunchKernelGGL(k_colfeat, dim3((unsigned)nblocks), dim3(256), 0, c->stream, d_items, nitems, d_base, d_out);
  return hipGetLastError() == hipSuccess ? RSG_OK : RSG_FAIL;
}

void rsg_feat_finish(const unsigned long long *cnt, int64_t N, int64_t L, const int *left, const int *right,
                     const double *ld, const double *rd, float *f)
{
  const int64_t m = N - 1;                      // internal nodes
  // branch statistics
  double T = 0., T2 = 0., term = 0., inter = 0.;
  int64_t nterm = 0, ninter = 0;
  for (int64_t v = 0; v < m; v++)
    for (int side = 0; side < 2; side++) {
      const double b = side ? rd[v] : ld[v];
      const int d = side ? right[v] : left[v];
      T += b; T2 += b * b;
      if (d > 0) { inter += b; ninter++; } else { term += b; nterm++; }
    }
  const double nb = (double)(2 * m), mb = T / nb;
  const double sdb = sqrt(fmax(0., T2 / nb - mb * mb));
  // depths (parents are reached before children in a walk from the root), subtree leaf counts
  std::vector<double> depth((size_t)m, 0.);
  std::vector<int> edges((size_t)m, 0), order;
  std::vector<int64_t> nleaf((size_t)m, 0);
  order.reserve((size_t)m);
  std::vector<int> stack(1, 0);
  while (!stack.empty()) {
    const int v = stack.back(); stack.pop_back();
    order.push_back(v);
    for (int side = 0; side < 2; side++) {
      const int d = side ? right[v] : left[v];
      if (d > 0) { depth[(size_t)d] = depth[(size_t)v] + (side ? rd[v] : ld[v]); edges[(size_t)d] = edges[(size_t)v] + 1; stack.push_back(d); }
    }
  }
  double sd = 0., sd2 = 0., dmax = 0., se = 0.;
  int64_t cherries = 0;
  for (int64_t v = 0; v < m; v++) {
    int leaves_here = 0;
    for (int side = 0; side < 2; side++) {
      const int d = side ? right[v] : left[v];
      if (d <= 0) {
        const double dl = depth[(size_t)v] + (side ? rd[v] : ld[v]);
        sd += dl; sd2 += dl * dl; dmax = fmax(dmax, dl); se += edges[(size_t)v] + 1;
        leaves_here++;
      }
    }
    if (leaves_here == 2) cherries++;
  }
  double colless = 0.;
  for (size_t k = order.size(); k-- > 0;) {           // children before parents
    const int v = order[k];
    const int64_t a = left[v] > 0 ? nleaf[(size_t)left[v]] : 1, b = right[v] > 0 ? nleaf[(size_t)right[v]] : 1;
    nleaf[(size_t)v] = a + b;
    colless += (double)(a > b ? a - b : b - a);
  }
  // split spectrum (nleaf is complete: a node's count follows its children's above)
  double s1[8] = {0., 0., 0., 0., 0., 0., 0., 0.}, s2[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  for (int64_t v = 0; v < m; v++)
    for (int side = 0; side < 2; side++) {
      const int d = side ? right[v] : left[v];
      const double t = side ? rd[v] : ld[v];
      const int64_t n = d > 0 ? nleaf[(size_t)d] : 1, nmin = n < N - n ? n : N - n;
      int k = 0;
      while (k < 7 && (nmin << (k + 2)) <= N) k++;
      const double e = 1. - exp(-t);
      s1[k] += t; s2[k] += e * e;
    }
  const double Nd = (double)N, md = sd / Nd;
  const double sdd = sqrt(fmax(0., sd2 / Nd - md * md));
  f[0] = (float)log(Nd); f[1] = (float)log((double)L); f[2] = (float)log(fmax(T, 1e-300));
  for (int k = 0; k < 16; k++) { f[3 + k] = (float)((double)cnt[k] / (double)L); f[19 + k] = (float)((double)cnt[16 + k] / (double)L); }
  const double res = (double)(cnt[32] + cnt[33] + cnt[34] + cnt[35]);
  for (int x = 0; x < 3; x++) f[35 + x] = (float)(res > 0. ? (double)cnt[32 + x] / res : 0.25);
  f[38] = (float)mb;
  f[39] = (float)(mb > 0. ? sdb / mb : 0.);
  f[40] = (float)md;
  f[41] = (float)(md > 0. ? sdd / md : 0.);
  f[42] = (float)(md > 0. ? dmax / md : 0.);
  f[43] = (float)(mb > 0. && nterm ? (term / (double)nterm) / mb : 0.);
  f[44] = (float)(mb > 0. && ninter ? (inter / (double)ninter) / mb : 0.);
  f[45] = (float)((se / Nd) / log2(Nd));
  f[46] = (float)((double)cherries / (Nd / 2.));
  f[47] = (float)(N > 2 ? colless / ((Nd - 1.) * (Nd - 2.) / 2.) : 0.);
  for (int k = 0; k < 8; k++) { f[48 + k] = (float)log(1e-3 + s1[k]); f[56 + k] = (float)log(1e-4 + s2[k]); }
}

extern "C" int rsg_null_features(rsg_ctx *c, rsg_msa *msa, const int *left, const int *right, const double *ld,
                                 const double *rd, float *feat, char *errbuf)
{
  if (!c || !msa || !left || !right || !ld || !rd || !feat) return rsg_fail(errbuf, "rsg_null_features: bad arguments");
  if (msa->N < 2) return rsg_fail(errbuf, "rsg_null_features: needs >= 2 sequences");
  HIPCHK(hipSetDevice(c->device));
  const int64_t nblk = (msa->L + 3) / 4;
  const int64_t item[6] = {0, msa->N, msa->Npad, msa->L, 0, msa->Lpad};
  int64_t *d_item = (int64_t *)c->d_small;                      // 48 B of context scratch
  unsigned long long *d_out = (unsigned long long *)c->d_small + 8;
  unsigned long long *h = (unsigned long long *)c->h_pinned;   // 36 of the 64 pinned doubles
  unsigned long long cnt[36];
  int st = RSG_OK;
  int64_t *h_item = (int64_t *)(c->h_pinned + 40);            // pinned slots 40..45: a fast upload
  memcpy(h_item, item, sizeof(item));
  if (hipMemcpyAsync(d_item, h_item, sizeof(item), hipMemcpyHostToDevice, c->stream) != hipSuccess) st = rsg_fail(errbuf, "upload failed");
  int t; rsg_tstart(c, "features", &t);
  if (st == RSG_OK && rsg_feat_launch(c, d_item, 1, nblk, msa->d_xt, d_out) != RSG_OK) st = rsg_fail(errbuf, "feature launch failed");
  rsg_tstop(c, t);
  if (st == RSG_OK && (hipMemcpyAsync(h, d_out, sizeof(cnt), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                       hipStreamSynchronize(c->stream) != hipSuccess)) st = rsg_fail(errbuf, "feature download failed");
  if (st == RSG_OK) memcpy(cnt, h, sizeof(cnt));
  if (st == RSG_OK) rsg_feat_finish(cnt, msa->N, msa->L, left, right, ld, rd, feat);
  return st;
}
