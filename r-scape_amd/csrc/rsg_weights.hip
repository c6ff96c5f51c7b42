// rsg_weights.hip -- sequence weights and the gap-column filter (the msaweight step of
// src/R-scape.c:1544-1562 and msamanip_RemoveGapColumns, src/msamanip.c:461-541).
//
// Digital RNA codes (Easel): 0..3 ACGU, 4 gap, 5..14 degenerate, 15 N, 16 nonresidue '*',
// 17 missing '~'.
//
// PB  (Henikoff & Henikoff 1994): column residue-type counts and the per-sequence sums are
//     device kernels in FP64; the normalisation to sum N is a one-block FP64 reduction.
// GSC (Gerstein, Sonnhammer & Chothia 1994): pairwise identities are a popcount contraction
//     over 2-bit packed sequences on the device (no MFMA: this is bit work); the UPGMA tree on
//     D = 1 - idents/min(len_a, len_b) and the top-down weight split run on the host in FP64.
#include "rsg_internal.h"
#include <cmath>
#include <cstring>
#include <vector>

namespace {

// ---------------- PB ----------------
// 64 columns x 4 sequence slices per block; inv[c*4+x] = 1/(ntypes * n_x) (0 if absent).
__global__ void __launch_bounds__(256) k_pb_cols(const uint8_t *__restrict__ ax, int64_t N, int64_t L,
                                                 double *__restrict__ inv)
{
  __shared__ int cnt[4][64][4];
  const int cl = threadIdx.x & 63, sl = threadIdx.x >> 6;
  const int64_t c = (int64_t)blockIdx.x * 64 + cl;
  int n[4] = {0, 0, 0, 0};
  if (c < L)
    for (int64_t s = sl; s < N; s += 4) {
      uint8_t a = ax[s * L + c];
      if (a < 4) n[a]++;
    }
#pragma unroll
  for (int x = 0; x < 4; x++) cnt[sl][cl][x] = n[x];
  __syncthreads();
  if (sl == 0 && c < L) {
    int t[4], ntypes = 0;
#pragma unroll
    for (int x = 0; x < 4; x++) { t[x] = cnt[0][cl][x] + cnt[1][cl][x] + cnt[2][cl][x] + cnt[3][cl][x]; ntypes += t[x] > 0; }
#pragma unroll
    for (int x = 0; x < 4; x++) inv[c * 4 + x] = t[x] ? 1.0 / (double)(ntypes * t[x]) : 0.0;
  }
}

template <typename T>
__device__ T block_sum256(T v, T *sh)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  T r = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return r;
}

// one block per sequence: w_s = sum_c inv[c][a_sc] / (#canonical residues of s)
__global__ void __launch_bounds__(256) k_pb_rows(const uint8_t *__restrict__ ax, int64_t L,
                                                 const double *__restrict__ inv, double *__restrict__ w)
{
  __shared__ double shd[4];
  __shared__ int64_t shi[4];
  const int64_t s = blockIdx.x;
  double acc = 0.;
  int64_t len = 0;
  for (int64_t c = threadIdx.x; c < L; c += 256) {
    uint8_t a = ax[s * L + c];
    if (a < 4) { acc += inv[c * 4 + a]; len++; }
  }
  acc = block_sum256(acc, shd);
  len = block_sum256(len, shi);
  if (threadIdx.x == 0) w[s] = len > 0 ? acc / (double)len : acc;
}

// one block: w *= N / sum(w), or all ones when the sum is 0
__global__ void __launch_bounds__(256) k_wnorm(double *__restrict__ w, int64_t N)
{
  __shared__ double shd[4];
  double acc = 0.;
  for (int64_t s = threadIdx.x; s < N; s += 256) acc += w[s];
  double sum = block_sum256(acc, shd);
  for (int64_t s = threadIdx.x; s < N; s += 256) w[s] = sum > 0. ? w[s] * ((double)N / sum) : 1.0;
}

// ---------------- GSC identities ----------------
// Sequence s, word k (32 columns): {lo bits, hi bits, canonical mask, 0}.
__global__ void k_pack2(const uint8_t *__restrict__ ax, int64_t N, int64_t L, int64_t W, int64_t Npad,
                        uint4 *__restrict__ P)
{
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Npad * W) return;
  int64_t s = t / W, k = t % W;
  uint32_t lo = 0, hi = 0, m = 0;
  if (s < N)
    for (int b = 0; b < 32; b++) {
      int64_t c = k * 32 + b;
      if (c >= L) break;
      uint8_t a = ax[s * L + c];
      if (a < 4) { m |= 1u << b; lo |= (uint32_t)(a & 1) << b; hi |= (uint32_t)(a >> 1) << b; }
    }
  P[s * W + k] = make_uint4(lo, hi, m, 0u);
}

// 64 x 64 sequence tile per block, 4 x 4 pairs per thread, words staged through LDS.
#define GSC_KW 16
__global__ void __launch_bounds__(256) k_ident(const uint4 *__restrict__ P, int64_t N, int64_t W,
                                               const int2 *__restrict__ tiles, int *__restrict__ id)
{
  __shared__ uint4 A[64][GSC_KW + 1], B[64][GSC_KW + 1];
  const int2 tl = tiles[blockIdx.x];
  const int64_t a0 = (int64_t)tl.x * 64, b0 = (int64_t)tl.y * 64;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  int acc[4][4] = {};
  for (int64_t k0 = 0; k0 < W; k0 += GSC_KW) {
    for (int e = threadIdx.x; e < 64 * GSC_KW; e += 256) {
      int r = e / GSC_KW, k = e % GSC_KW;
      uint4 z = make_uint4(0, 0, 0, 0);
      A[r][k] = (k0 + k < W) ? P[(a0 + r) * W + k0 + k] : z;
      B[r][k] = (k0 + k < W) ? P[(b0 + r) * W + k0 + k] : z;
    }
    __syncthreads();
    for (int k = 0; k < GSC_KW; k++) {
      uint4 va[4], vb[4];
#pragma unroll
      for (int i = 0; i < 4; i++) { va[i] = A[ty * 4 + i][k]; vb[i] = B[tx * 4 + i][k]; }
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
          acc[i][j] += __popc(va[i].z & vb[j].z & ~((va[i].x ^ vb[j].x) | (va[i].y ^ vb[j].y)));
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int64_t a = a0 + ty * 4 + i, b = b0 + tx * 4 + j;
      if (a < N && b < N) { id[a * N + b] = acc[i][j]; id[b * N + a] = acc[i][j]; }
    }
}

__global__ void k_lens(const uint4 *__restrict__ P, int64_t N, int64_t W, int *__restrict__ len)
{
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  int n = 0;
  for (int64_t k = 0; k < W; k++) n += __popc(P[s * W + k].z);
  len[s] = n;
}

struct DevBuf {
  std::vector<void *> p;
  ~DevBuf() { for (void *q : p) hipFree(q); }
  template <typename T> hipError_t get(T **out, size_t bytes) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes ? bytes : 1);
    if (e == hipSuccess) { p.push_back(q); *out = (T *)q; }
    return e;
  }
};

// UPGMA over the fractional-difference matrix D (N x N, destroyed).  Each step joins the
// first minimal active pair (r < c) in row-major order; the joined cluster takes row c and
// the last active row moves into row r.  Row minima are cached so a step costs O(n) plus
// the rows whose cached minimum moved.  Tree layout as ESL_TREE: internal nodes 0..N-2.
void upgma(std::vector<double> &D, int N, int *left, int *right, double *ld, double *rd)
{
  std::vector<int> idx(N), nin(N), amin(N);
  std::vector<double> vmin(N), height(N > 1 ? N - 1 : 1, 0.);
  for (int k = 0; k < N; k++) { idx[k] = -k; nin[k] = 1; }
  int n = N;
  auto rescan = [&](int r) {
    double m = INFINITY; int am = -1;
    const double *row = &D[(size_t)r * N];
    for (int c = r + 1; c < n; c++) if (row[c] < m) { m = row[c]; am = c; }
    vmin[r] = m; amin[r] = am;
  };
  for (int r = 0; r < n; r++) rescan(r);
  while (n >= 2) {
    int i = -1; double m = INFINITY;
    for (int r = 0; r + 1 < n; r++) if (vmin[r] < m) { m = vmin[r]; i = r; }
    if (i < 0) { i = 0; m = D[1]; }                  // every entry infinite: cannot happen for D in [0,1]
    int j = amin[i] < 0 ? i + 1 : amin[i];
    const int v = n - 2;
    left[v] = idx[i]; right[v] = idx[j];
    height[v] = m / 2.;
    ld[v] = rd[v] = height[v];
    if (idx[i] > 0) { ld[v] -= height[idx[i]]; if (ld[v] < 0.) ld[v] = 0.; }
    if (idx[j] > 0) { rd[v] -= height[idx[j]]; if (rd[v] < 0.) rd[v] = 0.; }
    for (int c = 0; c < n; c++) {
      if (c == i || c == j) continue;
      double x = (nin[j] * D[(size_t)j * N + c] + nin[i] * D[(size_t)i * N + c]) / (double)(nin[i] + nin[j]);
      D[(size_t)j * N + c] = x; D[(size_t)c * N + j] = x;
    }
    idx[j] = v; nin[j] += nin[i];
    const int last = n - 1;
    for (int c = 0; c < n; c++) { D[(size_t)i * N + c] = D[(size_t)last * N + c]; D[(size_t)c * N + i] = D[(size_t)c * N + last]; }
    idx[i] = idx[last]; nin[i] = nin[last];
    n--;
    // refresh the row-minimum cache for the active rows 0..n-1
    for (int r = 0; r + 1 < n; r++) {
      if (r == i || r == j || amin[r] == i || amin[r] == j || amin[r] == last || amin[r] >= n) { rescan(r); continue; }
      for (int c : {i, j}) {
        if (c <= r || c >= n) continue;
        double x = D[(size_t)r * N + c];
        if (x < vmin[r] || (x == vmin[r] && c < amin[r])) { vmin[r] = x; amin[r] = c; }
      }
    }
  }
}

} // namespace

extern "C" int rsg_weights_PB(rsg_ctx *c, const uint8_t *ax, int64_t N, int64_t L, double *wgt, char *errbuf)
{
  if (!c || !ax || !wgt || N < 1 || L < 1) return rsg_fail(errbuf, "rsg_weights_PB: bad arguments");
  HIPCHK(hipSetDevice(c->device));
  DevBuf db;
  uint8_t *d_ax; double *d_inv, *d_w;
  HIPCHK(db.get(&d_ax, (size_t)(N * L)));
  HIPCHK(db.get(&d_inv, sizeof(double) * 4 * (size_t)L));
  HIPCHK(db.get(&d_w, sizeof(double) * (size_t)N));
  HIPCHK(hipMemcpyAsync(d_ax, ax, (size_t)(N * L), hipMemcpyHostToDevice, c->stream));
  int t; rsg_tstart(c, "weights_pb", &t);
  hipLaunchKernelGGL(k_pb_cols, dim3((unsigned)((L + 63) / 64)), dim3(256), 0, c->stream, d_ax, N, L, d_inv);
  hipLaunchKernelGGL(k_pb_rows, dim3((unsigned)N), dim3(256), 0, c->stream, d_ax, L, d_inv, d_w);
  hipLaunchKernelGGL(k_wnorm, dim3(1), dim3(256), 0, c->stream, d_w, N);
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(wgt, d_w, sizeof(double) * (size_t)N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return RSG_OK;
}

extern "C" int rsg_weights_GSC(rsg_ctx *c, const uint8_t *ax, int64_t N, int64_t L, double *wgt, char *errbuf)
{
  if (!c || !ax || !wgt || N < 1 || L < 1) return rsg_fail(errbuf, "rsg_weights_GSC: bad arguments");
  if (N > 46340) return rsg_fail(errbuf, "rsg_weights_GSC: %lld sequences exceed the N x N identity table", (long long)N);
  if (N == 1) { wgt[0] = 1.0; return RSG_OK; }
  HIPCHK(hipSetDevice(c->device));
  const int64_t W = (L + 31) / 32, Npad = (N + 63) / 64 * 64, T = Npad / 64;
  DevBuf db;
  uint8_t *d_ax; uint4 *d_P; int *d_id, *d_len; int2 *d_tiles;
  std::vector<int2> tiles;
  for (int a = 0; a < T; a++) for (int b = a; b < T; b++) tiles.push_back(make_int2(a, b));
  HIPCHK(db.get(&d_ax, (size_t)(N * L)));
  HIPCHK(db.get(&d_P, sizeof(uint4) * (size_t)(Npad * W)));
  HIPCHK(db.get(&d_id, sizeof(int) * (size_t)(N * N)));
  HIPCHK(db.get(&d_len, sizeof(int) * (size_t)N));
  HIPCHK(db.get(&d_tiles, sizeof(int2) * tiles.size()));
  HIPCHK(hipMemcpyAsync(d_ax, ax, (size_t)(N * L), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice, c->stream));
  int t; rsg_tstart(c, "weights_gsc_ident", &t);
  hipLaunchKernelGGL(k_pack2, dim3((unsigned)((Npad * W + 255) / 256)), dim3(256), 0, c->stream, d_ax, N, L, W, Npad, d_P);
  hipLaunchKernelGGL(k_ident, dim3((unsigned)tiles.size()), dim3(256), 0, c->stream, d_P, N, W, d_tiles, d_id);
  hipLaunchKernelGGL(k_lens, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c->stream, d_P, N, W, d_len);
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  std::vector<int> id((size_t)(N * N)), len((size_t)N);
  HIPCHK(hipMemcpyAsync(id.data(), d_id, sizeof(int) * id.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(len.data(), d_len, sizeof(int) * len.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));

  const int n = (int)N;
  std::vector<double> D((size_t)n * n);
  for (int a = 0; a < n; a++) {
    D[(size_t)a * n + a] = 0.;
    for (int b = a + 1; b < n; b++) {
      int m = len[b] < len[a] ? len[b] : len[a];
      double pid = m == 0 ? 0. : (double)id[(size_t)a * n + b] / (double)m;
      D[(size_t)a * n + b] = D[(size_t)b * n + a] = 1. - pid;
    }
  }
  std::vector<int> left(n), right(n);
  std::vector<double> ld(n), rd(n), x(n);
  upgma(D, n, left.data(), right.data(), ld.data(), rd.data());
  // x[v]: total branch length below internal node v (children have larger indices)
  for (int v = n - 2; v >= 0; v--) {
    x[v] = ld[v] + rd[v];
    if (left[v] > 0) x[v] += x[left[v]];
    if (right[v] > 0) x[v] += x[right[v]];
  }
  // top-down: the weight carried into node v splits in proportion to each side's total length
  for (int v = 0; v <= n - 2; v++) {
    if (v == 0) x[v] = 0.;
    double lw = ld[v] + (left[v] > 0 ? x[left[v]] : 0.), rw = rd[v] + (right[v] > 0 ? x[right[v]] : 0.);
    double lx, rx;
    if (lw + rw == 0.) lx = rx = 0.5 * x[v];
    else { lx = x[v] * lw / (lw + rw); rx = x[v] * rw / (lw + rw); }
    if (left[v] <= 0) wgt[-left[v]] = lx + ld[v]; else x[left[v]] = lx + ld[v];
    if (right[v] <= 0) wgt[-right[v]] = rx + rd[v]; else x[right[v]] = rx + rd[v];
  }
  double s = 0., comp = 0.;                      // compensated sum
  for (int k = 0; k < n; k++) { double y = wgt[k] - comp, tt = s + y; comp = (tt - s) - y; s = tt; }
  for (int k = 0; k < n; k++) wgt[k] = s != 0. ? wgt[k] / s * (double)n : 1.0;
  return RSG_OK;
}

extern "C" int rsg_gapcol_mask(const uint8_t *ax, const double *wgt, int64_t N, int64_t L, double gapthresh, int *useme)
{
  if (!ax || !useme || N < 0 || L < 0) return -1;
  const double need = 1.0 - gapthresh;
  std::vector<double> res((size_t)L, 0.), tot((size_t)L, 0.);
  for (int64_t s = 0; s < N; s++) {
    const double w = wgt ? wgt[s] : 1.0;
    const uint8_t *row = ax + s * L;
    for (int64_t col = 0; col < L; col++) {
      uint8_t a = row[col];
      if (a < 4 || (a >= 5 && a <= 15)) { res[col] += w; tot[col] += w; }   // canonical, degenerate, N
      else if (a == 4) tot[col] += w;                                         // gap; '*' and '~' do not count
    }
  }
  int kept = 0;
  for (int64_t col = 0; col < L; col++) {
    useme[col] = (res[col] > 0. && res[col] / tot[col] >= need) ? 1 : 0;
    kept += useme[col];
  }
  return kept;
}

// ---------------- gap-column filter on the resident alignment ----------------
// msamanip_RemoveGapColumns (src/msamanip.c:461-541) on the device: one thread per column sums
// the weights of its residues (canonical, degenerate, N) and of residues + gaps in sequence
// order (the host loop's order, so the FP64 sums and the decisions are the same); nonresidue
// and missing symbols do not count.
__global__ void __launch_bounds__(64) k_gapcol(const uint8_t *__restrict__ ax, const double *__restrict__ w, int64_t N,
                                               int64_t L, double need, int *__restrict__ useme)
{
  const int64_t col = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (col >= L) return;
  double r = 0., tot = 0.;
  for (int64_t s = 0; s < N; s++) {
    const uint8_t a = ax[s * L + col];
    const double ws = w[s];
    if (a < 4 || (a >= 5 && a <= 15)) { r += ws; tot += ws; }
    else if (a == 4) tot += ws;
  }
  useme[col] = (r > 0. && r / tot >= need) ? 1 : 0;
}

// out (N x Lk) = the kept columns map[0..Lk) of ax (N x L)
__global__ void __launch_bounds__(256) k_colsubset(const uint8_t *__restrict__ ax, int64_t N, int64_t L, const int *__restrict__ map,
                                                   int64_t Lk, uint8_t *__restrict__ out)
{
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= N * Lk) return;
  const int64_t s = e / Lk, c = e - s * Lk;
  out[e] = ax[s * L + map[c]];
}

extern "C" int rsg_msa_RemoveGapColumns(rsg_ctx *c, const rsg_msa *src, double gapthresh, int *useme, int *map,
                                        rsg_msa **ret, char *errbuf)
{
  if (!c || !src || !ret) return rsg_fail(errbuf, "rsg_msa_RemoveGapColumns: bad arguments");
  *ret = nullptr;
  HIPCHK(hipSetDevice(c->device));
  const int64_t N = src->N, L = src->L;
  std::vector<int> use((size_t)L, 1);
  if (gapthresh < 1.0) {                                 // dofilter (useme is not an input here)
    DevBuf db;
    int *d_use;
    HIPCHK(db.get(&d_use, sizeof(int) * (size_t)L));
    int t; rsg_tstart(c, "gapcol", &t);
    hipLaunchKernelGGL(k_gapcol, dim3((unsigned)((L + 63) / 64)), dim3(64), 0, c->stream, src->d_ax, src->d_w, N, L,
                       1.0 - gapthresh, d_use);
    rsg_tstop(c, t);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(use.data(), d_use, sizeof(int) * (size_t)L, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  std::vector<int> mp;
  for (int64_t k = 0; k < L; k++) if (use[(size_t)k]) mp.push_back((int)k);
  if (useme) memcpy(useme, use.data(), sizeof(int) * (size_t)L);
  if (mp.empty()) return rsg_fail(errbuf, "no positions left after gap trimming");
  if (map) memcpy(map, mp.data(), sizeof(int) * mp.size());
  const int64_t Lk = (int64_t)mp.size();
  rsg_msa *m = nullptr;
  if (rsg_msa_alloc(c, N, Lk, &m, errbuf) != RSG_OK) return RSG_FAIL;
  {
    DevBuf db;
    int *d_map;
    if (db.get(&d_map, sizeof(int) * (size_t)Lk) != hipSuccess ||
        hipMemcpyAsync(d_map, mp.data(), sizeof(int) * (size_t)Lk, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
      rsg_msa_destroy(c, m);
      return rsg_fail(errbuf, "rsg_msa_RemoveGapColumns: column map upload failed");
    }
    hipLaunchKernelGGL(k_colsubset, dim3((unsigned)((N * Lk + 255) / 256)), dim3(256), 0, c->stream, src->d_ax, N, L, d_map,
                       Lk, m->d_ax);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(m->d_w, src->d_w, sizeof(double) * (size_t)N, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
        rsg_pack(c, m, errbuf) != RSG_OK || hipStreamSynchronize(c->stream) != hipSuccess) {
      rsg_msa_destroy(c, m);
      return rsg_fail(errbuf, "rsg_msa_RemoveGapColumns: column subset failed");
    }
  }
  m->weighted = src->weighted; m->wsum = src->wsum; m->wsum_q = src->wsum_q; m->weights_ok = src->weights_ok;
  *ret = m;
  return RSG_OK;
}
