// rsg_fitch.hip -- R-scape's own null alignment on the device: Fitch ancestral states,
// one shuffle of all columns, per-branch substitution replay.
//
// Reference: null_rscape (src/R-scape.c:1615-1724) -> Tree_FitchAlgorithmAncenstral
// (src/msatree.c:173-229, 1700-1931) -> msamanip_ShuffleTreeSubstitutions
// (src/msamanip.c:1449-1530, 1597-1780).  Symbols 0..4 (gap = 4 is a state); a leaf with
// another code enters the Fitch pass as one random state.  Draws are Philox4x32-10 keyed by
// (seed, purpose stream, node, position, tag) in place of Easel's sequential generator, so
// results match the reference in distribution and the oracle symbol for symbol.
//
// Rows: leaf n -> row n, internal node v -> row N + v (the reference's allmsa order).
// Work: the Fitch passes and the replay run one launch per tree depth (children before
// parents upward, parents before children downward); a column of a level is one thread in
// the Fitch passes, a branch is one wave in the replay (ballot compaction of the parent's
// positions holding each old symbol, then a partial Fisher-Yates on the positions that
// receive substitutions).
#include "rsg_internal.h"
#include <algorithm>
#include <vector>

namespace {

__device__ __forceinline__ int row_of(int d, int64_t N) { return d > 0 ? (int)(N + d) : -d; }

// Replicate r = blockIdx.y of every kernel: its seed is seeds[r]; its state sets and states
// (mask, all: 2N - 1 rows) start at r * rs, its shuffled leaf rows at r * N * L of `leaves` and
// its shuffled internal rows at r * (N - 1) * L of `inner`.

// leaf state sets: 1 << code for codes 0..4, else one random state
__global__ void k_fitch_leaves(const uint8_t *__restrict__ ax, int64_t N, int64_t L, const uint64_t *__restrict__ seeds,
                               uint8_t *__restrict__ mask, uint8_t *__restrict__ all, int64_t rs)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N * L) return;
  const uint64_t seed = seeds[blockIdx.y];
  mask += blockIdx.y * rs; all += blockIdx.y * rs;
  const int64_t n = e / L, c = e - n * L;
  const uint8_t a = ax[e];
  int s = a;
  if (a > 4) s = (int)(((uint64_t)rsg_philox(seed, 1u, (uint32_t)n, (uint32_t)c, 1u) * 5u) >> 32);
  mask[e] = (uint8_t)(1u << s);
  all[e] = a;
}

// upward pass over the internal nodes of one depth: S_v = S_l & S_r, or the residues of
// S_l | S_r when the intersection is empty
__global__ void k_fitch_up(const int *__restrict__ nodes, int nn, int64_t N, int64_t L, const int *__restrict__ left,
                           const int *__restrict__ right, uint8_t *__restrict__ mask, int64_t rs)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nn * L) return;
  mask += blockIdx.y * rs;
  const int v = nodes[e / L];
  const int64_t c = e % L;
  const unsigned sl = mask[(int64_t)row_of(left[v], N) * L + c], sr = mask[(int64_t)row_of(right[v], N) * L + c];
  const unsigned in = sl & sr;
  mask[(N + v) * L + c] = (uint8_t)(in ? in : ((sl | sr) & 0xFu));
}

// downward pass: keep the parent's state when the set holds it, then a uniform member
__global__ void k_fitch_down(const int *__restrict__ nodes, int nn, int64_t N, int64_t L, const int *__restrict__ parent,
                             const uint64_t *__restrict__ seeds, const uint8_t *__restrict__ mask,
                             uint8_t *__restrict__ all, int64_t rs)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nn * L) return;
  const uint64_t seed = seeds[blockIdx.y];
  mask += blockIdx.y * rs; all += blockIdx.y * rs;
  const int v = nodes[e / L];
  const int64_t c = e % L;
  unsigned S = mask[(N + v) * L + c];
  if (v > 0) {
    const unsigned pa = all[(N + parent[v]) * L + c];
    if ((S >> pa) & 1u) S = 1u << pa;
  }
  const unsigned m = __popc(S);
  const uint32_t u = rsg_philox(seed, 2u, (uint32_t)v, (uint32_t)c, 2u);
  unsigned idx = (unsigned)(((uint64_t)u * m) >> 32), st = 0;
  for (unsigned k = 0; k < 5; k++)
    if ((S >> k) & 1u) { if (idx == 0) { st = k; break; } idx--; }
  all[(N + v) * L + c] = (uint8_t)st;
}

// substitution counts of every branch b = 2 v + side: nsub[b][old * 5 + new], old != new
// (one wave per branch; a 25-bin LDS histogram, exact integer counts)
__global__ void __launch_bounds__(64) k_fitch_subs(int64_t N, int64_t L, const int *__restrict__ left,
                                                   const int *__restrict__ right, const uint8_t *__restrict__ all,
                                                   int *__restrict__ nsub, int64_t rs)
{
  __shared__ int hist[25];
  const int b = blockIdx.x, v = b >> 1, side = b & 1;
  all += blockIdx.y * rs;
  nsub += (int64_t)blockIdx.y * gridDim.x * 25;
  if (threadIdx.x < 25) hist[threadIdx.x] = 0;
  __syncthreads();
  const int d = side ? right[v] : left[v];
  const uint8_t *pa = all + (N + v) * L, *ch = all + (int64_t)row_of(d, N) * L;
  for (int64_t c = threadIdx.x; c < L; c += 64) {
    const int a = pa[c], x = ch[c];
    if (a != x && a <= 4 && x <= 4) atomicAdd(&hist[a * 5 + x], 1);
  }
  __syncthreads();
  if (threadIdx.x < 25) nsub[b * 25 + threadIdx.x] = hist[threadIdx.x];
}

// the shuffled root: column c of the root takes column perm[c] of its ancestral states
__global__ void k_gather_root(const uint8_t *__restrict__ all, const int *__restrict__ perm, int64_t N, int64_t L,
                              uint8_t *__restrict__ inner, int64_t rs)
{
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= L) return;
  all += blockIdx.y * rs;
  perm += blockIdx.y * L;
  inner += blockIdx.y * (N - 1) * L;
  inner[c] = all[N * L + perm[c]];
}

// Orders a wave's global stores before its later loads (lanes exchange data through memory;
// the waves of a block work on different branches and never wait for each other).
__device__ __forceinline__ void wave_sync()
{
  __threadfence_block();
  __builtin_amdgcn_wave_barrier();
}

// Replay of the branches below the internal nodes of one depth: one wave per branch.  For each
// old symbol: the parent's positions holding it (ascending, ballot compaction), a partial
// Fisher-Yates whose swap indices the 64 lanes draw ahead (a draw depends only on its step, not
// on the array) and lane 0 applies in order, then the positions that receive substitutions
// written in parallel.  LCOLS: the position list in LDS (L <= FITCH_LDS_L), else in cols_all
// (L ints per wave slot, 2 maxw slots per replicate).
#define FITCH_LDS_L 4000
template <bool LCOLS>
__global__ void __launch_bounds__(256) k_fitch_replay(const int *__restrict__ nodes, int nn, int64_t N, int64_t L,
                                                      const int *__restrict__ left, const int *__restrict__ right,
                                                      const uint64_t *__restrict__ seeds, const int *__restrict__ nsub,
                                                      uint8_t *__restrict__ leaves, uint8_t *__restrict__ inner,
                                                      int *__restrict__ cols_all, int maxw)
{
  extern __shared__ int lds_cols[];                           // LCOLS: 4 x L ints
  __shared__ int kbuf[4][64];                                 // drawn swap indices, per wave
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int slot = blockIdx.x * 4 + wv;                       // branch slot in this level
  if (slot >= 2 * nn) return;
  const int64_t r = blockIdx.y;
  const uint64_t seed = seeds[r];
  nsub += r * 2 * (N - 1) * 25;
  leaves += r * N * L;
  inner += r * (N - 1) * L;
  const int v = nodes[slot >> 1], side = slot & 1;
  const int d = side ? right[v] : left[v];
  const int chrow = row_of(d, N);
  const int b = 2 * v + side;
  const uint8_t *pa = inner + (int64_t)v * L;
  uint8_t *ch = (d > 0) ? inner + (int64_t)d * L : leaves + (int64_t)(-d) * L;
  int *cols = LCOLS ? lds_cols + wv * L : cols_all + (r * 2 * maxw + slot) * L;
  for (int64_t c = lane; c < L; c += 64) ch[c] = pa[c];
  wave_sync();
  for (int oldc = 0; oldc < 5; oldc++) {
    int cnt[5], nst = 0;
    for (int nw = 0; nw < 5; nw++) { cnt[nw] = nsub[b * 25 + oldc * 5 + nw]; nst += cnt[nw]; }
    if (!nst) continue;
    // positions of the parent holding oldc, ascending (ballot compaction)
    int ncol = 0;
    for (int64_t c0 = 0; c0 < L; c0 += 64) {
      const int64_t c = c0 + lane;
      const bool hit = c < L && pa[c] == oldc;
      const unsigned long long bal = __ballot(hit);
      if (hit) cols[ncol + __popcll(bal & ((1ull << lane) - 1ull))] = (int)c;
      ncol += __popcll(bal);
    }
    wave_sync();                                              // cols visible to lane 0
    if (ncol == 0) continue;
    // partial Fisher-Yates: step n (ncol down to nlow) swaps positions n - 1 and k_n; only the
    // last take = min(nst, ncol) positions receive substitutions
    const int take = nst < ncol ? nst : ncol;
    const int nlow = (ncol - take + 1 > 2) ? ncol - take + 1 : 2;
    for (int n0 = ncol; n0 >= nlow; n0 -= 64) {
      const int n = n0 - lane;
      if (n >= nlow) {
        const uint32_t u = rsg_philox(seed, 4u, (uint32_t)(chrow * 5 + oldc), (uint32_t)n, 4u);
        kbuf[wv][lane] = (int)(((uint64_t)u * (uint64_t)n) >> 32);
      }
      wave_sync();
      if (lane == 0) {
        const int steps = (n0 - nlow + 1 < 64) ? n0 - nlow + 1 : 64;
        for (int t = 0; t < steps; t++) {
          const int nn1 = n0 - t - 1, k = kbuf[wv][t];
          const int x = cols[nn1]; cols[nn1] = cols[k]; cols[k] = x;
        }
      }
      wave_sync();
    }
    // position ncol - 1 - t (t < take) gets the new symbol whose cumulative count first exceeds t
    for (int t = lane; t < take; t += 64) {
      int acc = 0, sym = 4;
      for (int nw = 0; nw < 5; nw++) { if (t < acc + cnt[nw]) { sym = nw; break; } acc += cnt[nw]; }
      ch[cols[ncol - 1 - t]] = (uint8_t)sym;
    }
    wave_sync();
  }
}

struct Buf {
  std::vector<void *> p;
  ~Buf() { for (void *q : p) hipFree(q); }
  template <typename T> bool get(T **o, size_t n)
  {
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return false;
    p.push_back(q); *o = (T *)q;
    return true;
  }
};

} // namespace

int rsg_fitch_plan(int64_t N, int64_t L, const int *left, const int *right, FitchPlan &P, char *errbuf)
{
  if (!left || !right || N < 2 || L < 1) return rsg_fail(errbuf, "Fitch null: bad arguments");
  if (N > (1 << 28) || L > (1 << 30)) return rsg_fail(errbuf, "Fitch null: alignment too large");
  const int m = (int)(N - 1);
  // depths of the internal nodes (walk from the root; malformed trees rejected), parents
  std::vector<int> depth((size_t)m, -1), stack(1, 0);
  P.parent.assign((size_t)m, 0);
  depth[0] = 0;
  int seen = 1, maxd = 0;
  while (!stack.empty()) {
    const int v = stack.back(); stack.pop_back();
    for (int side = 0; side < 2; side++) {
      const int d = side ? right[v] : left[v];
      if (d > 0) {
        if (d >= m || depth[(size_t)d] >= 0) return rsg_fail(errbuf, "Fitch null: malformed tree at node %d", v);
        depth[(size_t)d] = depth[(size_t)v] + 1; P.parent[(size_t)d] = v; maxd = std::max(maxd, depth[(size_t)d]);
        stack.push_back(d); seen++;
      } else if (-d >= N) return rsg_fail(errbuf, "Fitch null: leaf index %d out of range", -d);
    }
  }
  if (seen != m) return rsg_fail(errbuf, "Fitch null: %d internal nodes unreachable from the root", m - seen);
  P.lev0.assign((size_t)maxd + 2, 0);
  for (int v = 0; v < m; v++) P.lev0[(size_t)depth[(size_t)v] + 1]++;
  for (int k = 0; k <= maxd; k++) P.lev0[(size_t)k + 1] += P.lev0[(size_t)k];
  P.nodes.resize((size_t)m);
  {
    std::vector<int> fill(P.lev0.begin(), P.lev0.end() - 1);
    for (int v = 0; v < m; v++) P.nodes[(size_t)fill[(size_t)depth[(size_t)v]]++] = v;
  }
  P.maxw = 0;
  for (int k = 0; k <= maxd; k++) P.maxw = std::max(P.maxw, P.lev0[(size_t)k + 1] - P.lev0[(size_t)k]);
  P.N = N; P.L = L; P.m = m; P.maxd = maxd;
  return RSG_OK;
}

// one permutation of all columns (Fisher-Yates as esl_vec_IShuffle)
void rsg_fitch_perm(uint64_t seed, int64_t L, int *perm)
{
  for (int64_t k = 0; k < L; k++) perm[k] = (int)k;
  for (int64_t n = L; n > 1; n--) {
    const uint32_t u = rsg_philox(seed, 3u, 0u, (uint32_t)n, 3u);
    const int64_t k = (int64_t)(((uint64_t)u * (uint64_t)n) >> 32);
    std::swap(perm[n - 1], perm[k]);
  }
}

// Fitch ancestral states only (Tree_FitchAlgorithmAncenstral): D.all receives the leaves' codes
// and every internal node's state; D.mask, D.all, D.ax, the tree arrays and D.seeds are used.
int rsg_fitch_ancestral(rsg_ctx *c, const FitchPlan &P, const FitchDev &D, int R)
{
  const int64_t N = P.N, L = P.L, rs = (2 * N - 1) * L;
  const unsigned ry = (unsigned)R;
  hipLaunchKernelGGL(k_fitch_leaves, dim3((unsigned)((N * L + 255) / 256), ry), dim3(256), 0, c->stream, D.ax, N, L,
                     D.seeds, D.mask, D.all, rs);
  for (int k = P.maxd; k >= 0; k--) {
    const int nn = P.lev0[(size_t)k + 1] - P.lev0[(size_t)k];
    hipLaunchKernelGGL(k_fitch_up, dim3((unsigned)(((int64_t)nn * L + 255) / 256), ry), dim3(256), 0, c->stream,
                       D.nodes + P.lev0[(size_t)k], nn, N, L, D.left, D.right, D.mask, rs);
  }
  for (int k = 0; k <= P.maxd; k++) {
    const int nn = P.lev0[(size_t)k + 1] - P.lev0[(size_t)k];
    hipLaunchKernelGGL(k_fitch_down, dim3((unsigned)(((int64_t)nn * L + 255) / 256), ry), dim3(256), 0, c->stream,
                       D.nodes + P.lev0[(size_t)k], nn, N, L, D.parent, D.seeds, D.mask, D.all, rs);
  }
  return hipGetLastError() == hipSuccess ? RSG_OK : RSG_FAIL;
}

int rsg_fitch_launch(rsg_ctx *c, const FitchPlan &P, const FitchDev &D, int R)
{
  const int64_t N = P.N, L = P.L, rs = (2 * N - 1) * L;
  const unsigned ry = (unsigned)R;
  hipLaunchKernelGGL(k_fitch_leaves, dim3((unsigned)((N * L + 255) / 256), ry), dim3(256), 0, c->stream, D.ax, N, L,
                     D.seeds, D.mask, D.all, rs);
  for (int k = P.maxd; k >= 0; k--) {
    const int nn = P.lev0[(size_t)k + 1] - P.lev0[(size_t)k];
    hipLaunchKernelGGL(k_fitch_up, dim3((unsigned)(((int64_t)nn * L + 255) / 256), ry), dim3(256), 0, c->stream,
                       D.nodes + P.lev0[(size_t)k], nn, N, L, D.left, D.right, D.mask, rs);
  }
  for (int k = 0; k <= P.maxd; k++) {
    const int nn = P.lev0[(size_t)k + 1] - P.lev0[(size_t)k];
    hipLaunchKernelGGL(k_fitch_down, dim3((unsigned)(((int64_t)nn * L + 255) / 256), ry), dim3(256), 0, c->stream,
                       D.nodes + P.lev0[(size_t)k], nn, N, L, D.parent, D.seeds, D.mask, D.all, rs);
  }
  hipLaunchKernelGGL(k_fitch_subs, dim3((unsigned)(2 * P.m), ry), dim3(64), 0, c->stream, N, L, D.left, D.right, D.all,
                     D.nsub, rs);
  hipLaunchKernelGGL(k_gather_root, dim3((unsigned)((L + 255) / 256), ry), dim3(256), 0, c->stream, D.all, D.perm, N, L,
                     D.inner, rs);
  const bool lcols = L <= FITCH_LDS_L;
  for (int k = 0; k <= P.maxd; k++) {
    const int nn = P.lev0[(size_t)k + 1] - P.lev0[(size_t)k];
    const dim3 grid((unsigned)((2 * nn + 3) / 4), ry);
    if (lcols)
      hipLaunchKernelGGL(k_fitch_replay<true>, grid, dim3(256), (size_t)(4 * L) * sizeof(int), c->stream,
                         D.nodes + P.lev0[(size_t)k], nn, N, L, D.left, D.right, D.seeds, D.nsub, D.leaves, D.inner,
                         D.cols, P.maxw);
    else
      hipLaunchKernelGGL(k_fitch_replay<false>, grid, dim3(256), 0, c->stream,
                         D.nodes + P.lev0[(size_t)k], nn, N, L, D.left, D.right, D.seeds, D.nsub, D.leaves, D.inner,
                         D.cols, P.maxw);
  }
  return hipGetLastError() == hipSuccess ? RSG_OK : RSG_FAIL;
}

extern "C" int rsg_null_fitch(rsg_ctx *c, const uint8_t *ax, int64_t N, int64_t L, const int *left, const int *right,
                              uint64_t seed, uint8_t *leaves, char *errbuf)
{
  if (!c || !ax || !left || !right || !leaves) return rsg_fail(errbuf, "rsg_null_fitch: bad arguments");
  HIPCHK(hipSetDevice(c->device));
  FitchPlan P;
  if (rsg_fitch_plan(N, L, left, right, P, errbuf) != RSG_OK) return RSG_FAIL;
  std::vector<int> perm((size_t)L);
  rsg_fitch_perm(seed, L, perm.data());
  const int64_t nrow = 2 * N - 1, m = P.m;
  Buf B;
  uint8_t *d_ax, *d_mask, *d_all, *d_inner, *d_leaves; int *d_left, *d_right, *d_parent, *d_nodes, *d_perm, *d_nsub, *d_cols;
  uint64_t *d_seed;
  if (!B.get(&d_ax, (size_t)(N * L)) || !B.get(&d_mask, (size_t)(nrow * L)) || !B.get(&d_all, (size_t)(nrow * L)) ||
      !B.get(&d_inner, (size_t)((N - 1) * L)) || !B.get(&d_leaves, (size_t)(N * L)) || !B.get(&d_left, (size_t)m) ||
      !B.get(&d_right, (size_t)m) || !B.get(&d_parent, (size_t)m) || !B.get(&d_nodes, (size_t)m) ||
      !B.get(&d_perm, (size_t)L) || !B.get(&d_nsub, (size_t)(2 * m) * 25) || !B.get(&d_cols, (size_t)(2 * P.maxw) * (size_t)L) ||
      !B.get(&d_seed, 1))
    return rsg_fail(errbuf, "rsg_null_fitch: out of device memory");
  HIPCHK(hipMemcpyAsync(d_ax, ax, (size_t)(N * L), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_left, left, sizeof(int) * m, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_right, right, sizeof(int) * m, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_parent, P.parent.data(), sizeof(int) * m, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_nodes, P.nodes.data(), sizeof(int) * m, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_perm, perm.data(), sizeof(int) * L, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_seed, &seed, sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
  FitchDev D{d_ax, d_left, d_right, d_parent, d_nodes, d_perm, d_seed, d_mask, d_all, d_inner, d_leaves, d_nsub, d_cols};
  int t; rsg_tstart(c, "fitch_null", &t);
  const int st = rsg_fitch_launch(c, P, D, 1);
  rsg_tstop(c, t);
  if (st != RSG_OK) return rsg_fail(errbuf, "rsg_null_fitch: launch failed");
  HIPCHK(hipMemcpyAsync(leaves, d_leaves, (size_t)(N * L), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return RSG_OK;
}
