// rsg_sim.hip -- phylogenetic null simulation on the device.
//
// Reference: cov_GenerateAlignment (src/cov_simulate.c:60-86) -> Ungapped :88-159 ->
// cov_evolve_root_ungapped_tree :288-324 -> per branch P(t) = exp(tQ)
// (src/ratematrix.c:185-233, esl_dmx_Exp) and inverse-CDF draws cov_addres / cov_addpair
// (:756-798); trees from esl_tree_Simulate rescaled to a mean branch length
// (src/R-scape-sim.c:608-680).
//
// Design: a tree is the Easel ESL_TREE layout (internal nodes 0..N-2, root 0, parents
// before children; children > 0 internal, <= 0 leaf -child).  Sequence rows: internal
// node v -> row v, leaf l -> row N-1+l.  Every symbol draw is Philox(seed, stream,
// row, column, tag) compared against uint32 CDF thresholds of the branch's P(t) row, so a
// draw depends only on the parent's symbol in the same column: the evolution runs one
// launch per tree depth (Yule trees are O(log N) deep), each covering every (node at that
// depth, column) of every alignment in the batch.  Internal rows live in HBM between levels;
// leaves go straight to the output.
// P(t) = Pi^-1/2 V diag(exp(lambda t)) V^T Pi^1/2 from the eigendecomposition of the
// symmetrised GTR generator (host, FP64 Jacobi); the per-branch tables are built on the
// device in FP64.
#include "rsg_internal.h"
#include <cmath>
#include <cstring>
#include <algorithm>
#include <climits>
#include <vector>
#include <atomic>
#include <thread>

// Transition tables are rounded as the host computes them (no FMA contraction).
#pragma clang fp contract(off)

// ---------------------------------------------------------------- host: model
// Eigendecomposition of a symmetric 4x4 (cyclic Jacobi).
static void jacobi4(double A[4][4], double V[4][4], double lam[4])
{
  for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) V[i][j] = (i == j);
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.;
    for (int p = 0; p < 4; p++) for (int q = p + 1; q < 4; q++) off += A[p][q] * A[p][q];
    if (off < 1e-30) break;
    for (int p = 0; p < 4; p++)
      for (int q = p + 1; q < 4; q++) {
        if (fabs(A[p][q]) < 1e-300) continue;
        double th = (A[q][q] - A[p][p]) / (2 * A[p][q]);
        double t = (th >= 0 ? 1. : -1.) / (fabs(th) + sqrt(th * th + 1.));
        double cs = 1. / sqrt(t * t + 1.), sn = t * cs;
        for (int k = 0; k < 4; k++) {
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = cs * akp - sn * akq; A[k][q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < 4; k++) {
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = cs * apk - sn * aqk; A[q][k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < 4; k++) {
          double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = cs * vkp - sn * vkq; V[k][q] = sn * vkp + cs * vkq;
        }
      }
  }
  for (int i = 0; i < 4; i++) lam[i] = A[i][i];
}

// Model packed for the device: [0..15] M1 = Pi^-1/2 V, [16..31] M2 = V^T Pi^1/2, [32..35] lambda,
// [36..39] pi.  P(t)[a][b] = sum_k M1[a][k] exp(lambda_k t) M2[k][b].
extern "C" int rsg_gtr_model(const double *pi, const double *exch, double *model)
{
  double Q[4][4] = {{0}};
  const int pr[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  for (int k = 0; k < 6; k++) {
    int a = pr[k][0], b = pr[k][1];
    Q[a][b] = exch[k] * pi[b]; Q[b][a] = exch[k] * pi[a];
  }
  double mu = 0.;
  for (int a = 0; a < 4; a++) {
    double s = 0.;
    for (int b = 0; b < 4; b++) if (b != a) s += Q[a][b];
    Q[a][a] = -s; mu += pi[a] * s;
  }
  if (!(mu > 0.)) return RSG_FAIL;
  for (int a = 0; a < 4; a++) for (int b = 0; b < 4; b++) Q[a][b] /= mu;   // one substitution per unit time
  double S[4][4], V[4][4], lam[4];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) S[a][b] = sqrt(pi[a]) * Q[a][b] / sqrt(pi[b]);
  jacobi4(S, V, lam);
  for (int a = 0; a < 4; a++)
    for (int k = 0; k < 4; k++) {
      model[a * 4 + k] = V[a][k] / sqrt(pi[a]);
      model[16 + k * 4 + a] = V[a][k] * sqrt(pi[a]);
    }
  for (int k = 0; k < 4; k++) { model[32 + k] = lam[k]; model[36 + k] = pi[k]; }
  return RSG_OK;
}

// Yule tree: lineages split one at a time (uniform choice), waiting times Exp(k); then the
// branch lengths are rescaled to mean branch length mean_bl.  Draws from Philox(seed, stream).
extern "C" int rsg_tree_yule(int64_t n, double mean_bl, uint64_t seed, uint32_t stream,
                             int *left, int *right, double *ld, double *rd)
{
  if (n < 2) return RSG_FAIL;
  auto unif = [&](uint32_t ctr, uint32_t tag) {
    return ((double)rsg_philox(seed, stream, 0xFFFFFFF0u, ctr, tag) + 0.5) / 4294967296.0;
  };
  std::vector<double> birth((size_t)(n - 1));
  std::vector<std::pair<int, int>> lin;            // (parent internal node, side)
  lin.reserve((size_t)n);
  lin.push_back(std::make_pair(-1, 0));
  double t = 0.;
  for (int64_t v = 0; v < n - 1; v++) {
    double k = (double)lin.size();
    t += -log(unif((uint32_t)v, 1u)) / k;
    size_t idx = (size_t)(unif((uint32_t)v, 2u) * k);
    if (idx >= lin.size()) idx = lin.size() - 1;
    std::pair<int, int> pl = lin[idx];
    lin[idx] = lin.back(); lin.pop_back();
    birth[(size_t)v] = t;
    if (pl.first >= 0) { if (pl.second == 0) left[pl.first] = (int)v; else right[pl.first] = (int)v; }
    lin.push_back(std::make_pair((int)v, 0));
    lin.push_back(std::make_pair((int)v, 1));
  }
  double t_end = t - log(unif((uint32_t)n, 3u)) / (double)lin.size();
  int leaf = 0;
  for (auto &pl : lin) {
    if (pl.second == 0) left[pl.first] = -leaf; else right[pl.first] = -leaf;
    leaf++;
  }
  double sum = 0.;
  for (int64_t v = 0; v < n - 1; v++) {
    double tl = (left[v] > 0) ? birth[(size_t)left[v]] : t_end;
    double tr = (right[v] > 0) ? birth[(size_t)right[v]] : t_end;
    ld[v] = tl - birth[(size_t)v]; rd[v] = tr - birth[(size_t)v];
    sum += ld[v] + rd[v];
  }
  double scale = mean_bl / (sum / (2. * (double)(n - 1)));
  for (int64_t v = 0; v < n - 1; v++) { ld[v] *= scale; rd[v] *= scale; }
  return RSG_OK;
}

// ---------------------------------------------------------------- device: tables + evolution
// CDF thresholds of one P(t) row: c_k = floor(2^32 * sum_{b<=k} P[a][b]) for k = 0..2.
__device__ void row_cdf(const double *model, double t, int a, uint32_t *out)
{
  double e[4];
  for (int k = 0; k < 4; k++) e[k] = exp(model[32 + k] * t);
  double cum = 0.;
  for (int b = 0; b < 3; b++) {
    double p = 0.;
    for (int k = 0; k < 4; k++) p += model[a * 4 + k] * e[k] * model[16 + k * 4 + b];
    cum += (p > 0.) ? p : 0.;
    double v = cum * 4294967296.0;
    out[b] = (v >= 4294967295.0) ? 0xFFFFFFFFu : (uint32_t)v;
  }
}

__global__ void k_cdf_list(const double *__restrict__ model, const double *__restrict__ t, int64_t nt,
                           uint32_t *__restrict__ cdf)
{
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nt * 4) return;
  row_cdf(model, t[idx / 4], (int)(idx % 4), cdf + (idx / 4) * 12 + (idx % 4) * 3);
}

extern "C" int rsg_gtr_cdf(rsg_ctx *c, const double *model, int64_t nt, const double *t, uint32_t *cdf, char *errbuf)
{
  double *d_m = nullptr, *d_t = nullptr; uint32_t *d_c = nullptr;
  HIPCHK(hipMalloc(&d_m, sizeof(double) * 40));
  HIPCHK(hipMalloc(&d_t, sizeof(double) * (size_t)std::max<int64_t>(nt, 1)));
  HIPCHK(hipMalloc(&d_c, sizeof(uint32_t) * 12 * (size_t)std::max<int64_t>(nt, 1)));
  HIPCHK(hipMemcpy(d_m, model, sizeof(double) * 40, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_t, t, sizeof(double) * (size_t)nt, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_cdf_list, dim3((unsigned)((nt * 4 + 255) / 256)), dim3(256), 0, c->stream, d_m, d_t, nt, d_c);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(cdf, d_c, sizeof(uint32_t) * 12 * (size_t)nt, hipMemcpyDeviceToHost));
  hipFree(d_m); hipFree(d_t); hipFree(d_c);
  return RSG_OK;
}

__device__ __forceinline__ int icdf3(uint32_t u, const uint32_t *c)
{
  return (u >= c[0]) + (u >= c[1]) + (u >= c[2]);
}

__device__ __forceinline__ int find_sim_item(const SimItem *__restrict__ items, int nitems, int64_t blk)
{
  int lo = 0, hi = nitems - 1;
  while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (items[mid].col_blk0 <= blk) lo = mid; else hi = mid - 1; }
  return lo;
}

// Root row of every item from the stationary distribution; one thread per column.
__global__ void __launch_bounds__(64)
k_root(const SimItem *__restrict__ items, int nitems, const double *__restrict__ models, uint8_t *__restrict__ seqs)
{
  const SimItem it = items[find_sim_item(items, nitems, blockIdx.x)];
  const int64_t col = (blockIdx.x - it.col_blk0) * 64 + threadIdx.x;
  if (col >= it.L) return;
  const double *pi = models + it.model_off + 36;
  const uint32_t u = rsg_philox(it.seed, it.stream, (uint32_t)(2 * it.N - 1), (uint32_t)col, 0u);
  double cum = 0.; int a = 3;
  for (int b = 0; b < 3; b++) {
    cum += pi[b];
    const double v = cum * 4294967296.0;
    const uint32_t th = (v >= 4294967295.0) ? 0xFFFFFFFFu : (uint32_t)v;
    if (u < th) { a = b; break; }
  }
  seqs[it.seq_off + col] = (uint8_t)a;
}

// One tree level: every internal node v at this depth draws its two children.  Wave w of block
// b takes entry 4 b + w of the level (ent[k] = {item, v, -, -}, ne entries) and walks the item's
// columns in steps of 64 (one lane per column; every wave has work, and the node's item, table
// and children are read once per wave).
__global__ void __launch_bounds__(256)
k_evolve_level(const SimItem *__restrict__ items, const int4 *__restrict__ ent, int64_t ne,
               const int *__restrict__ left, const int *__restrict__ right, const uint32_t *__restrict__ cdf,
               uint8_t *__restrict__ seqs, uint8_t *__restrict__ leaves)
{
  const int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= ne) return;
  const int lane = threadIdx.x & 63;
  const int4 e = ent[k];
  const SimItem it = items[e.x];
  const int64_t L = it.L, v = e.y;
  uint8_t *S = seqs + it.seq_off;
  const uint32_t *cd = cdf + (it.tree_off + v) * 24;
  const int dc[2] = {left[it.tree_off + v], right[it.tree_off + v]};
  for (int64_t col = lane; col < L; col += 64) {
    const int par = S[v * L + col];
    uint32_t w[4];                                     // both children's draws, one block
    rsg_philox4(it.seed, it.stream, (uint32_t)v, (uint32_t)col, RSG_TAG_CHILDREN, w);
#pragma unroll
    for (int side = 0; side < 2; side++) {
      const int d = dc[side];
      int sym = icdf3(w[side], cd + side * 12 + par * 3);
      if (d > 0) S[(int64_t)d * L + col] = (uint8_t)sym;
      else {
        if (it.gapthr && w[2 + side] < it.gapthr) sym = 4;
        leaves[it.leaf_off + (int64_t)(-d) * L + col] = (uint8_t)sym;
      }
    }
  }
}

// Per-item branch tables into the [node][side][4][3] layout used by k_evolve_level.
__global__ void k_branch_tables(const SimItem *__restrict__ items, int nitems, const double *__restrict__ ld,
                                const double *__restrict__ rd, const double *__restrict__ models,
                                int64_t total_nodes, uint32_t *__restrict__ cdf)
{
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;     // (node, side, row)
  if (idx >= total_nodes * 8) return;
  const int64_t node = idx / 8; const int side = (int)((idx / 4) & 1); const int a = (int)(idx & 3);
  int lo = 0, hi = nitems - 1;
  while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (items[mid].tree_off <= node) lo = mid; else hi = mid - 1; }
  const double t = side ? rd[node] : ld[node];
  row_cdf(models + items[lo].model_off, t, a, cdf + node * 24 + side * 12 + a * 3);
}

// Tree levels of all items (host).  Depths come from a walk down from each root, so any
// node order is accepted; a node reached twice or never is an error.
// Depth of every internal node of one tree (0 = root), checking that the tree is well formed:
// every internal node reached exactly once from the root, leaves in range.  Trees whose
// children always carry larger indices than their parent (as Yule simulation and UPGMA number
// them) take one forward pass; any other tree a depth-first walk.
static int tree_depths(const int *left, const int *right, int64_t N, int *depth, std::vector<int> &stack, int item,
                       char *errbuf)
{
  const int64_t m = N - 1;
  for (int64_t v = 0; v < m; v++) depth[v] = -1;
  depth[0] = 0;
  int64_t seen = 1;
  bool forward = true;
  for (int64_t v = 0; v < m && forward; v++) {
    if (depth[v] < 0) { forward = false; break; }          // not reached yet: not in topological order
    for (int side = 0; side < 2; side++) {
      const int d = side ? right[v] : left[v];
      if (d > 0) {
        if (d >= m || depth[d] >= 0) return rsg_fail(errbuf, "item %d: malformed tree at node %lld", item, (long long)v);
        if (d <= v) { forward = false; break; }
        depth[d] = depth[v] + 1; seen++;
      } else if (-d >= N) return rsg_fail(errbuf, "item %d: leaf index %d out of range", item, -d);
    }
  }
  if (!forward) {
    for (int64_t v = 0; v < m; v++) depth[v] = -1;
    depth[0] = 0; seen = 1;
    stack.assign(1, 0);
    while (!stack.empty()) {
      const int v = stack.back(); stack.pop_back();
      for (int side = 0; side < 2; side++) {
        const int d = side ? right[v] : left[v];
        if (d > 0) {
          if (d >= m || depth[d] >= 0) return rsg_fail(errbuf, "item %d: malformed tree at node %d", item, v);
          depth[d] = depth[v] + 1; stack.push_back(d); seen++;
        } else if (-d >= N) return rsg_fail(errbuf, "item %d: leaf index %d out of range", item, -d);
      }
    }
  }
  if (seen != m) return rsg_fail(errbuf, "item %d: %lld internal nodes unreachable from the root", item, (long long)(m - seen));
  return RSG_OK;
}

int rsg_sim_levels(const SimItem *items, int nitems, const int *left, const int *right, SimLevels &lv, char *errbuf)
{
  // Items in contiguous ranges of about equal node counts, one per host thread.  Pass 1: depths
  // of the range's nodes and its node / 64-column block counts per depth.  Thread 0 then merges
  // the counts (level offsets, and each range's start within every level: the ranges in item
  // order), and pass 2 places the entries.  A node's draws are keyed by its item and row, so the
  // order of a level's entries only sets which blocks simulate it: items ascending, each item's
  // nodes ascending, as one thread would place them.
  int64_t nodes = 0;
  for (int k = 0; k < nitems; k++) nodes += items[k].N - 1;
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)rsg_host_threads(), (int64_t)nitems, nodes / 16384 + 1}));
  std::vector<int> rb((size_t)nt + 1, nitems);
  rb[0] = 0;
  {
    int64_t acc = 0;
    int t = 1;
    for (int k = 0; k < nitems && t < nt; k++) {
      acc += items[k].N - 1;
      while (t < nt && acc * nt >= nodes * t) rb[(size_t)t++] = k + 1;
    }
  }
  struct Range { std::vector<int64_t> cnt, blk, mxl, fill, bfill; int st = RSG_OK; char err[RSG_ERRBUFSIZE]; };
  std::vector<Range> R((size_t)nt);
  std::vector<int> depth((size_t)nodes);
  lv.ent.alloc((size_t)nodes);
  std::atomic<int> arrived{0}, merged{0};
  int status = RSG_OK;
  auto work = [&](int t) {
    Range &r = R[(size_t)t];
    std::vector<int> stack;
    for (int k = rb[(size_t)t]; k < rb[(size_t)t + 1] && r.st == RSG_OK; k++) {
      const SimItem &it = items[k];
      int *dk = depth.data() + it.tree_off;
      if (tree_depths(left + it.tree_off, right + it.tree_off, it.N, dk, stack, k, r.err) != RSG_OK) { r.st = RSG_FAIL; break; }
      int dmax = 0;
      for (int64_t v = 0; v < it.N - 1; v++) dmax = std::max(dmax, dk[v]);
      if ((int64_t)r.cnt.size() <= dmax) { r.cnt.resize((size_t)dmax + 1, 0); r.blk.resize((size_t)dmax + 1, 0); r.mxl.resize((size_t)dmax + 1, 0); }
      const int64_t nbk = (it.L + 63) / 64;
      for (int64_t v = 0; v < it.N - 1; v++) {
        const size_t dv = (size_t)dk[v];
        r.cnt[dv]++; r.blk[dv] += nbk;
        if (it.L > r.mxl[dv]) r.mxl[dv] = it.L;
      }
    }
    arrived.fetch_add(1);
    if (t == 0) {
      while (arrived.load() < nt) std::this_thread::yield();
      for (int u = 0; u < nt && status == RSG_OK; u++)
        if (R[(size_t)u].st != RSG_OK) status = rsg_fail(errbuf, "%s", R[(size_t)u].err);
      size_t nlev = 0;
      for (const Range &q : R) nlev = std::max(nlev, q.cnt.size());
      if (status == RSG_OK) {
        lv.lev_ent0.assign(nlev + 1, 0);
        lv.lev_blocks.assign(nlev, 0);
        lv.lev_maxl.assign(nlev, 0);
        for (const Range &q : R)
          for (size_t l = 0; l < q.cnt.size(); l++) {
            lv.lev_ent0[l + 1] += q.cnt[l]; lv.lev_blocks[l] += q.blk[l]; lv.lev_maxl[l] = std::max(lv.lev_maxl[l], q.mxl[l]);
          }
        for (size_t l = 0; l < nlev; l++) {
          if (lv.lev_blocks[l] > INT32_MAX) { status = rsg_fail(errbuf, "too many column blocks in one tree level"); break; }
          lv.lev_ent0[l + 1] += lv.lev_ent0[l];
        }
        std::vector<int64_t> f(lv.lev_ent0.begin(), lv.lev_ent0.end() - 1), bf(nlev, 0);
        for (Range &q : R) {
          q.fill = f; q.bfill = bf;
          for (size_t l = 0; l < q.cnt.size(); l++) { f[l] += q.cnt[l]; bf[l] += q.blk[l]; }
        }
      }
      merged.store(1);
    } else {
      while (!merged.load()) std::this_thread::yield();
    }
    if (status != RSG_OK) return;
    for (int k = rb[(size_t)t]; k < rb[(size_t)t + 1]; k++) {
      const SimItem &it = items[k];
      const int *dk = depth.data() + it.tree_off;
      const int64_t nbk = (it.L + 63) / 64;
      for (int64_t v = 0; v < it.N - 1; v++) {
        const size_t dv = (size_t)dk[v];
        lv.ent[(size_t)r.fill[dv]++] = make_int4(k, (int)v, (int)r.bfill[dv], 0);
        r.bfill[dv] += nbk;
      }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; t++) th.emplace_back(work, t);
  work(0);
  for (auto &x : th) x.join();
  return status;
}

// Launch-only entry (no allocation, no synchronisation; usable under stream capture).
int rsg_sim_launch(rsg_ctx *c, const SimItem *d_items, int nitems, int64_t nodes, int64_t blocks,
                   const SimLevels &lv, const int4 *d_ent, const int *d_left, const int *d_right,
                   const double *d_ld, const double *d_rd, const double *d_models, uint32_t *d_cdf,
                   uint8_t *d_seqs, uint8_t *d_leaves, const uint8_t *root0, int64_t root_len)
{
  if (nodes > 0)
    hipLaunchKernelGGL(k_branch_tables, dim3((unsigned)((nodes * 8 + 255) / 256)), dim3(256), 0, c->stream,
                       d_items, nitems, d_ld, d_rd, d_models, nodes, d_cdf);
  if (blocks > 0)
    hipLaunchKernelGGL(k_root, dim3((unsigned)blocks), dim3(64), 0, c->stream, d_items, nitems, d_models, d_seqs);
  // a given root row (cov_add_root, src/cov_simulate.c:204-238) replaces item 0's drawn one
  if (root0 && hipMemcpyAsync(d_seqs, root0, (size_t)root_len, hipMemcpyHostToDevice, c->stream) != hipSuccess) return RSG_FAIL;
  for (size_t l = 0; l < lv.lev_blocks.size(); l++) {
    const int64_t e0 = lv.lev_ent0[l], ne = lv.lev_ent0[l + 1] - e0;
    if (lv.lev_blocks[l] > 0)
      hipLaunchKernelGGL(k_evolve_level, dim3((unsigned)((ne + 3) / 4)), dim3(256), 0,
                         c->stream, d_items, d_ent + e0, ne, d_left, d_right, d_cdf, d_seqs, d_leaves);
  }
  return hipGetLastError() == hipSuccess ? RSG_OK : RSG_FAIL;
}

// ---------------------------------------------------------------- host API
namespace {
struct SimBuffers {
  std::vector<void *> p;
  template <typename T> bool get(T **out, size_t count)
  {
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return false;
    p.push_back(q); *out = (T *)q;
    return true;
  }
  ~SimBuffers() { for (void *q : p) hipFree(q); }
};
}

// Simulate a set of alignments into device leaves (row-major N x L per item at leaf offsets).
// Trees come from the caller (host arrays, concatenated per item); models = nitems x 40
// doubles of rsg_gtr_model output.
int rsg_simulate_device(rsg_ctx *c, int nitems, const int64_t *N, const int64_t *L,
                        const int *left, const int *right, const double *ld, const double *rd,
                        const double *models, const uint64_t *seed, const uint32_t *stream, const double *gapfrac,
                        uint8_t *d_leaves, const int64_t *leaf_off, char *errbuf, const uint8_t *root0)
{
  std::vector<SimItem> items((size_t)nitems);
  int64_t nodes = 0, seq = 0, blocks = 0;
  for (int k = 0; k < nitems; k++) {
    SimItem &it = items[(size_t)k];
    it.N = N[k]; it.L = L[k]; it.tree_off = nodes; it.seq_off = seq; it.leaf_off = leaf_off[k];
    it.seed = seed[k]; it.stream = stream[k]; it.model_off = 40 * (int64_t)k;
    const double g = gapfrac ? gapfrac[k] : 0.;
    it.gapthr = (g <= 0.) ? 0u : (g >= 1. ? 0xFFFFFFFFu : (uint32_t)(g * 4294967296.0));
    it.col_blk0 = blocks;
    nodes += N[k] - 1; seq += (N[k] - 1) * L[k]; blocks += (L[k] + 63) / 64;
  }
  SimLevels lv;
  if (rsg_sim_levels(items.data(), nitems, left, right, lv, errbuf) != RSG_OK) return RSG_FAIL;
  SimBuffers b;
  int *d_left, *d_right; double *d_ld, *d_rd, *d_models; uint32_t *d_cdf; uint8_t *d_seqs; SimItem *d_items; int4 *d_ent;
  if (!b.get(&d_left, (size_t)nodes) || !b.get(&d_right, (size_t)nodes) || !b.get(&d_ld, (size_t)nodes) ||
      !b.get(&d_rd, (size_t)nodes) || !b.get(&d_cdf, 24 * (size_t)nodes) || !b.get(&d_models, 40 * (size_t)nitems) ||
      !b.get(&d_seqs, (size_t)seq) || !b.get(&d_items, (size_t)nitems) || !b.get(&d_ent, lv.ent.size()))
    return rsg_fail(errbuf, "simulation buffers: out of device memory");
  HIPCHK(hipMemcpyAsync(d_left, left, sizeof(int) * nodes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_right, right, sizeof(int) * nodes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_ld, ld, sizeof(double) * nodes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_rd, rd, sizeof(double) * nodes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_models, models, sizeof(double) * 40 * nitems, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_items, items.data(), sizeof(SimItem) * nitems, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_ent, lv.ent.data(), sizeof(int4) * lv.ent.size(), hipMemcpyHostToDevice, c->stream));
  int t; rsg_tstart(c, "simulate", &t);
  const int lr = rsg_sim_launch(c, d_items, nitems, nodes, blocks, lv, d_ent, d_left, d_right, d_ld, d_rd, d_models,
                                d_cdf, d_seqs, d_leaves, root0, root0 ? L[0] : 0);
  rsg_tstop(c, t);
  if (lr != RSG_OK) return rsg_fail(errbuf, "simulation launch failed");
  HIPCHK(hipStreamSynchronize(c->stream));
  return RSG_OK;
}

// Simulate one alignment and return its leaves (N x L) to the host.
extern "C" int rsg_null_simulate(rsg_ctx *c, const rsg_null_job *job, uint8_t *leaves, char *errbuf)
{
  return rsg_sim_one(c, job, nullptr, leaves, errbuf);
}

int rsg_sim_one(rsg_ctx *c, const rsg_null_job *job, const uint8_t *root, uint8_t *leaves, char *errbuf)
{
  HIPCHK(hipSetDevice(c->device));
  if (job->nleaf < 2 || job->alen < 1) return rsg_fail(errbuf, "simulation needs >= 2 leaves and >= 1 column");
  uint8_t *d_leaves = nullptr;
  HIPCHK(hipMalloc(&d_leaves, (size_t)(job->nleaf * job->alen)));
  int64_t off = 0;
  int rc = rsg_simulate_device(c, 1, &job->nleaf, &job->alen, job->left, job->right, job->ld, job->rd,
                               job->model, &job->seed, &job->stream, &job->gapfrac, d_leaves, &off, errbuf, root);
  if (rc == RSG_OK) {
    if (hipMemcpy(leaves, d_leaves, (size_t)(job->nleaf * job->alen), hipMemcpyDeviceToHost) != hipSuccess)
      rc = rsg_fail(errbuf, "copy of simulated leaves failed");
  }
  hipFree(d_leaves);
  return rc;
}

// ---------------------------------------------------------------- structured nulls: pair model
// cov_GenerateAlignmentUngapped with a structure (noss off): a base-paired column pair (i, ct[i])
// evolves as one 16-state character under the ribosum pair rate matrix (src/cov_simulate.c:393-394,
// 406-412: P(t) = ratematrix_ConditionalsFromRate(t, bprsQ)); the root pair is drawn from the
// stationary pair frequencies bprsM and a child pair from its parent pair's P(t) row
// (cov_addpair, :774-798).  Unpaired columns evolve under the 4-state GTR model as before.
// Device order: the 4-state level kernel draws every column of a level's children, then the
// pair kernel overwrites both columns of every pair; draws are Philox keyed by (seed, stream,
// child row, 5' column, tag 2), leaf gaps reuse the leaf-gap words of the 4-state path's
// per-node blocks (internal node, column, RSG_TAG_CHILDREN).

// Eigendecomposition of a symmetric n x n (cyclic Jacobi; A, V row-major n x n).
static void jacobi_n(int n, double *A, double *V, double *lam)
{
  for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) V[i * n + j] = (i == j);
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.;
    for (int p = 0; p < n; p++) for (int q = p + 1; q < n; q++) off += A[p * n + q] * A[p * n + q];
    if (off < 1e-30) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        if (fabs(A[p * n + q]) < 1e-300) continue;
        double th = (A[q * n + q] - A[p * n + p]) / (2 * A[p * n + q]);
        double t = (th >= 0 ? 1. : -1.) / (fabs(th) + sqrt(th * th + 1.));
        double cs = 1. / sqrt(t * t + 1.), sn = t * cs;
        for (int k = 0; k < n; k++) {
          double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = cs * akp - sn * akq; A[k * n + q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = cs * apk - sn * aqk; A[q * n + k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < n; k++) {
          double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = cs * vkp - sn * vkq; V[k * n + q] = sn * vkp + cs * vkq;
        }
      }
  }
  for (int i = 0; i < n; i++) lam[i] = A[i * n + i];
}

// Packed: [0..255] Pi^-1/2 V, [256..511] V^T Pi^1/2, [512..527] eigenvalues, [528..543] M16.
// P(t)[a][b] = sum_k model[a * 16 + k] exp(lambda_k t) model[256 + k * 16 + b].
extern "C" int rsg_pair_model(const double *Q16, const double *M16, double *model)
{
  if (!Q16 || !M16 || !model) return RSG_FAIL;
  double msum = 0., qmax = 0.;
  for (int a = 0; a < 16; a++) {
    if (!(M16[a] > 0.)) return RSG_FAIL;                       // every pair state reachable
    msum += M16[a];
    double rs = 0.;
    for (int b = 0; b < 16; b++) {
      if (a != b && !(Q16[a * 16 + b] >= 0.)) return RSG_FAIL;
      rs += Q16[a * 16 + b]; qmax = std::max(qmax, fabs(Q16[a * 16 + b]));
    }
    if (fabs(rs) > 1e-9 * std::max(1., qmax)) return RSG_FAIL;   // rows of a rate matrix sum to 0
  }
  if (fabs(msum - 1.) > 1e-9) return RSG_FAIL;
  std::vector<double> S(256), V(256), lam(16);
  for (int a = 0; a < 16; a++)
    for (int b = 0; b < 16; b++) S[a * 16 + b] = sqrt(M16[a]) * Q16[a * 16 + b] / sqrt(M16[b]);
  for (int a = 0; a < 16; a++)                                  // reversible: M16[a] Q[a][b] = M16[b] Q[b][a]
    for (int b = a + 1; b < 16; b++) {
      if (fabs(S[a * 16 + b] - S[b * 16 + a]) > 1e-9 * std::max(1., qmax)) return RSG_FAIL;
      const double s = 0.5 * (S[a * 16 + b] + S[b * 16 + a]);
      S[a * 16 + b] = S[b * 16 + a] = s;
    }
  jacobi_n(16, S.data(), V.data(), lam.data());
  for (int a = 0; a < 16; a++)
    for (int k = 0; k < 16; k++) {
      model[a * 16 + k] = V[a * 16 + k] / sqrt(M16[a]);
      model[256 + k * 16 + a] = V[a * 16 + k] * sqrt(M16[a]);
    }
  for (int k = 0; k < 16; k++) { model[512 + k] = lam[(size_t)k]; model[528 + k] = M16[k]; }
  return RSG_OK;
}

// CDF thresholds of one 16-state P(t) row: c_k = floor(2^32 sum_{b<=k} P[a][b]), k = 0..14.
__device__ void row_cdf16(const double *m, double t, int a, uint32_t *out)
{
  double e[16];
  for (int k = 0; k < 16; k++) e[k] = exp(m[512 + k] * t);
  double cum = 0.;
  for (int b = 0; b < 15; b++) {
    double p = 0.;
    for (int k = 0; k < 16; k++) p += m[a * 16 + k] * e[k] * m[256 + k * 16 + b];
    cum += (p > 0.) ? p : 0.;
    const double v = cum * 4294967296.0;
    out[b] = (v >= 4294967295.0) ? 0xFFFFFFFFu : (uint32_t)v;
  }
}

__device__ __forceinline__ int icdf15(uint32_t u, const uint32_t *c)
{
  int k = 0;
#pragma unroll
  for (int b = 0; b < 15; b++) k += (u >= c[b]);
  return k;
}

// Pair tables [node][side][16 parent states][15]
__global__ void k_pair_tables(const double *__restrict__ m16, const double *__restrict__ ld, const double *__restrict__ rd,
                              int64_t nodes, uint32_t *__restrict__ tab)
{
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;     // (node, side, parent state)
  if (idx >= nodes * 32) return;
  const int64_t node = idx / 32; const int side = (int)((idx / 16) & 1), a = (int)(idx & 15);
  row_cdf16(m16, side ? rd[node] : ld[node], a, tab + (node * 2 + side) * 240 + a * 15);
}

// Root pairs from the stationary pair frequencies (key of row 2N - 1, 5' column, tag 2).
__global__ void __launch_bounds__(64)
k_pair_root(const double *__restrict__ m16, const int2 *__restrict__ pairs, int npairs, int64_t N, uint64_t seed,
            uint32_t stream, uint8_t *__restrict__ root)
{
  const int p = blockIdx.x * 64 + threadIdx.x;
  if (p >= npairs) return;
  const int2 pr = pairs[p];
  uint32_t th[15];
  double cum = 0.;
  for (int b = 0; b < 15; b++) {
    cum += m16[528 + b];
    const double v = cum * 4294967296.0;
    th[b] = (v >= 4294967295.0) ? 0xFFFFFFFFu : (uint32_t)v;
  }
  const int st = icdf15(rsg_philox(seed, stream, (uint32_t)(2 * N - 1), (uint32_t)pr.x, 2u), th);
  root[pr.x] = (uint8_t)(st >> 2);
  root[pr.y] = (uint8_t)(st & 3);
}

// One tree level: the pairs of both children of every node at this depth (blockIdx.y).
__global__ void __launch_bounds__(64)
k_pair_level(const int *__restrict__ nodes, const int2 *__restrict__ pairs, int npairs, const int *__restrict__ left,
             const int *__restrict__ right, const uint32_t *__restrict__ tab, int64_t N, int64_t L, uint64_t seed,
             uint32_t stream, uint32_t gapthr, uint8_t *__restrict__ S, uint8_t *__restrict__ leaves)
{
  const int p = blockIdx.x * 64 + threadIdx.x;
  if (p >= npairs) return;
  const int v = nodes[blockIdx.y];
  const int2 pr = pairs[p];
  const int a = 4 * S[(int64_t)v * L + pr.x] + S[(int64_t)v * L + pr.y];
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const int d = side ? right[v] : left[v];
    const int64_t row = (d > 0) ? d : (N - 1 - d);
    const int st = icdf15(rsg_philox(seed, stream, (uint32_t)row, (uint32_t)pr.x, 2u), tab + ((int64_t)v * 2 + side) * 240 + a * 15);
    int x = st >> 2, y = st & 3;
    if (d > 0) { S[(int64_t)d * L + pr.x] = (uint8_t)x; S[(int64_t)d * L + pr.y] = (uint8_t)y; }
    else {
      // the leaf-gap draws of the unpaired evolution of these columns, so gaps do not depend
      // on the pairing
      if (gapthr) {
        uint32_t wx[4], wy[4];
        rsg_philox4(seed, stream, (uint32_t)v, (uint32_t)pr.x, RSG_TAG_CHILDREN, wx);
        rsg_philox4(seed, stream, (uint32_t)v, (uint32_t)pr.y, RSG_TAG_CHILDREN, wy);
        if (wx[2 + side] < gapthr) x = 4;
        if (wy[2 + side] < gapthr) y = 4;
      }
      leaves[(int64_t)(-d) * L + pr.x] = (uint8_t)x;
      leaves[(int64_t)(-d) * L + pr.y] = (uint8_t)y;
    }
  }
}

extern "C" int rsg_null_simulate_pairs(rsg_ctx *c, const rsg_null_job *job, const int *ct, const double *model16,
                                       uint8_t *leaves, char *errbuf)
{
  return rsg_sim_pairs(c, job, ct, model16, nullptr, leaves, errbuf);
}

int rsg_sim_pairs(rsg_ctx *c, const rsg_null_job *job, const int *ct, const double *model16, const uint8_t *root,
                  uint8_t *leaves, char *errbuf)
{
  if (!c || !job || !ct || !model16 || !leaves) return rsg_fail(errbuf, "rsg_null_simulate_pairs: bad arguments");
  HIPCHK(hipSetDevice(c->device));
  const int64_t N = job->nleaf, L = job->alen;
  if (N < 2 || L < 1) return rsg_fail(errbuf, "simulation needs >= 2 leaves and >= 1 column");
  if (L > INT_MAX) return rsg_fail(errbuf, "alignment too long");
  std::vector<int2> pairs;
  for (int64_t i = 0; i < L; i++) {
    const int j = ct[i];
    if (j < 0) continue;
    if (j >= L || j == i || ct[j] != (int)i) return rsg_fail(errbuf, "rsg_null_simulate_pairs: ct is not a pairing at column %lld", (long long)i);
    if (j > i) pairs.push_back(make_int2((int)i, j));
  }
  SimItem it{};
  it.N = N; it.L = L; it.tree_off = 0; it.seq_off = 0; it.leaf_off = 0; it.seed = job->seed; it.stream = job->stream;
  it.gapthr = (job->gapfrac <= 0.) ? 0u : (job->gapfrac >= 1. ? 0xFFFFFFFFu : (uint32_t)(job->gapfrac * 4294967296.0));
  it.col_blk0 = 0; it.model_off = 0;
  SimLevels lv;
  if (rsg_sim_levels(&it, 1, job->left, job->right, lv, errbuf) != RSG_OK) return RSG_FAIL;
  const int64_t nodes = N - 1, blocks = (L + 63) / 64, np = (int64_t)pairs.size();
  std::vector<int> lev_nodes(lv.ent.size());
  for (size_t k = 0; k < lv.ent.size(); k++) lev_nodes[k] = lv.ent[k].y;
  SimBuffers b;
  int *d_left, *d_right, *d_nodes; double *d_ld, *d_rd, *d_models, *d_m16; uint32_t *d_cdf, *d_tab; uint8_t *d_seqs, *d_leaves;
  SimItem *d_items; int4 *d_ent; int2 *d_pairs;
  if (!b.get(&d_left, (size_t)nodes) || !b.get(&d_right, (size_t)nodes) || !b.get(&d_ld, (size_t)nodes) ||
      !b.get(&d_rd, (size_t)nodes) || !b.get(&d_cdf, 24 * (size_t)nodes) || !b.get(&d_models, 40) || !b.get(&d_m16, 544) ||
      !b.get(&d_tab, 480 * (size_t)nodes) || !b.get(&d_seqs, (size_t)(nodes * L)) || !b.get(&d_leaves, (size_t)(N * L)) ||
      !b.get(&d_items, 1) || !b.get(&d_ent, lv.ent.size()) || !b.get(&d_pairs, (size_t)np) || !b.get(&d_nodes, lev_nodes.size()))
    return rsg_fail(errbuf, "simulation buffers: out of device memory");
  HIPCHK(hipMemcpyAsync(d_left, job->left, sizeof(int) * nodes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_right, job->right, sizeof(int) * nodes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_ld, job->ld, sizeof(double) * nodes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_rd, job->rd, sizeof(double) * nodes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_models, job->model, sizeof(double) * 40, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_m16, model16, sizeof(double) * 544, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_items, &it, sizeof(SimItem), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_ent, lv.ent.data(), sizeof(int4) * lv.ent.size(), hipMemcpyHostToDevice, c->stream));
  if (np) HIPCHK(hipMemcpyAsync(d_pairs, pairs.data(), sizeof(int2) * np, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_nodes, lev_nodes.data(), sizeof(int) * lev_nodes.size(), hipMemcpyHostToDevice, c->stream));
  int t; rsg_tstart(c, "simulate_pairs", &t);
  hipLaunchKernelGGL(k_branch_tables, dim3((unsigned)((nodes * 8 + 255) / 256)), dim3(256), 0, c->stream,
                     d_items, 1, d_ld, d_rd, d_models, nodes, d_cdf);
  hipLaunchKernelGGL(k_pair_tables, dim3((unsigned)((nodes * 32 + 255) / 256)), dim3(256), 0, c->stream, d_m16, d_ld, d_rd,
                     nodes, d_tab);
  hipLaunchKernelGGL(k_root, dim3((unsigned)blocks), dim3(64), 0, c->stream, d_items, 1, d_models, d_seqs);
  if (np) hipLaunchKernelGGL(k_pair_root, dim3((unsigned)((np + 63) / 64)), dim3(64), 0, c->stream, d_m16, d_pairs, (int)np, N,
                             it.seed, it.stream, d_seqs);
  if (root) HIPCHK(hipMemcpyAsync(d_seqs, root, (size_t)L, hipMemcpyHostToDevice, c->stream));   // cov_add_root
  for (size_t l = 0; l < lv.lev_blocks.size(); l++) {
    const int64_t e0 = lv.lev_ent0[l], ne = lv.lev_ent0[l + 1] - e0;
    if (lv.lev_blocks[l] > 0)
      hipLaunchKernelGGL(k_evolve_level, dim3((unsigned)((ne + 3) / 4)), dim3(256), 0,
                         c->stream, d_items, d_ent + e0, ne, d_left, d_right, d_cdf, d_seqs, d_leaves);
    if (np && ne > 0)
      hipLaunchKernelGGL(k_pair_level, dim3((unsigned)((np + 63) / 64), (unsigned)ne), dim3(64), 0, c->stream,
                         d_nodes + e0, d_pairs, (int)np, d_left, d_right, d_tab, N, L, it.seed, it.stream, it.gapthr,
                         d_seqs, d_leaves);
  }
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(leaves, d_leaves, (size_t)(N * L), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return RSG_OK;
}
