// rsg_indel.hip -- cov_GenerateAlignment with a given root and the e1 indel model
// (src/cov_simulate.c:60-86: cov_GenerateAlignmentUngapped, then cov_GenerateAlignmentAddGaps,
// :161-190, 326-361, 495-563, 634-722, 803-859).
//
// The reference grows the alignment branch by branch in tree order: each branch walks an
// e1 HMM (states B, S, D, I) over the alignment as it is then (every column, also those earlier
// branches inserted), deleting the child's cell at D and inserting a column (gap in every other
// row) at I.  A branch's walk depends on the others only through the number of columns it walks,
// and its transitions do not depend on the position, so the walk over the first m columns is a
// prefix of the walk over any longer alignment.  Here every branch walks, in parallel on the
// device (one thread per branch, counter-based draws), an upper bound Lmax of columns and records
// its D and I events; the host then takes each branch's prefix by a scan over the branches in
// tree order (L_b+1 = L_b + insertions of branch b within its first L_b columns), places every
// inserted column in the final alignment by an order-statistics pass over the insertions in
// reverse (a Fenwick tree of free slots), maps each deletion to its final column by a forward pass
// (a Fenwick tree of the columns present at that moment), and the device builds the leaves by one
// gather from the ungapped leaves and one scatter of the leaf branches' deletions and inserted
// residues.  Draws: step k of the branch into row r is Philox (seed, stream, r, k, tag 3); the
// residue of its n-th insertion is (seed, stream, r, n, tag 4).
#include "rsg_internal.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

uint32_t th32(double x) { return (x >= 1.0) ? 0xFFFFFFFFu : (uint32_t)(x * 4294967296.0); }

// one thread per branch: the walk over Lmax columns; events (i << 1) | (1 for I, 0 for D), i the
// number of columns consumed when the event happens; cnt > cap means the buffer overflowed
__global__ void __launch_bounds__(64) k_indel_walk(int nb, const uint32_t *__restrict__ th, const int *__restrict__ rows,
                                                   uint64_t seed, uint32_t stream, int64_t Lmax, int cap,
                                                   int *__restrict__ ev, int *__restrict__ cnt)
{
  const int b = blockIdx.x * 64 + threadIdx.x;
  if (b >= nb) return;
  uint32_t t[8];
#pragma unroll
  for (int k = 0; k < 8; k++) t[k] = th[8 * b + k];
  const uint32_t row = (uint32_t)rows[b];
  int64_t i = 0;
  uint32_t step = 0;
  int st = 0, n = 0;                                   // B
  for (;;) {
    const uint32_t u = rsg_philox(seed, stream, row, step++, 3u);
    st = (u < t[2 * st]) ? 1 : (u < t[2 * st + 1]) ? 2 : 3;   // S, D, I
    if (st != 3) i++;
    if (i == Lmax + 1) break;
    if (st != 1) {
      if (n < cap) ev[(int64_t)b * cap + n] = (int)((i << 1) | (st == 3 ? 1 : 0));
      if (++n > cap) break;                            // overflow: the host retries with more room
    }
  }
  cnt[b] = n;
}

__global__ void __launch_bounds__(256) k_indel_gather(const uint8_t *__restrict__ ung, int64_t N, int64_t L0,
                                                      const int *__restrict__ src, int64_t F, uint8_t *__restrict__ out)
{
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= N * F) return;
  const int64_t l = e / F, c = e - l * F;
  const int s = src[c];
  out[e] = (s >= 0) ? ung[l * L0 + s] : (uint8_t)4;
}

// {leaf, final column, row, insertion ordinal (-1: a deletion)}
__global__ void __launch_bounds__(256) k_indel_scatter(const int4 *__restrict__ ent, int n, int64_t F, uint64_t seed,
                                                       uint32_t stream, const uint32_t *__restrict__ ith,
                                                       uint8_t *__restrict__ out)
{
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const int4 e = ent[k];
  uint8_t v = 4;
  if (e.w >= 0) {
    const uint32_t u = rsg_philox(seed, stream, (uint32_t)e.z, (uint32_t)e.w, 4u);
    v = (uint8_t)((u >= ith[0]) + (u >= ith[1]) + (u >= ith[2]));
  }
  out[(int64_t)e.x * F + e.y] = v;
}

struct Fenwick {
  std::vector<int> t;
  int n = 0, lg = 1;
  explicit Fenwick(int n_) : t((size_t)n_ + 1, 0), n(n_) { while ((1 << lg) <= n) lg++; }
  void add(int i, int v) { for (i++; i <= n; i += i & -i) t[(size_t)i] += v; }
  int kth(int k) const                               // 0-based index of the k-th (1-based) one
  {
    int pos = 0;
    for (int s = lg; s >= 0; s--) {
      const int nx = pos + (1 << s);
      if (nx <= n && t[(size_t)nx] < k) { pos = nx; k -= t[(size_t)nx]; }
    }
    return pos;
  }
};

struct DevMem {
  std::vector<void *> p;
  ~DevMem() { for (void *q : p) hipFree(q); }
  template <typename T> bool get(T **out, size_t count)
  {
    void *q = nullptr;
    if (hipMalloc(&q, sizeof(T) * (count ? count : 1)) != hipSuccess) return false;
    p.push_back(q); *out = (T *)q; return true;
  }
};

} // namespace

extern "C" int rsg_e1_transitions(const rsg_e1_rates *R, double time, int64_t L, double *t)
{
  if (!R || !t || !(time >= 0.)) return RSG_FAIL;
  const double gM = 1.0 - std::exp(-R->muAM * time), gD = 1.0 - std::exp(-R->muAD * time), gI = 1.0 - std::exp(-R->muAI * time);
  double beta = 0.0;
  if (time != 0.0 && R->ldI != 0.0) {
    // the TKF insertion term of e1_model_LI_BetaFunc (src/e1_model.c:634-690)
    const double ld = R->ldI, mu = R->muI, a = ld - mu, expo = -time * std::fabs(a);
    double num, den;
    if (std::fabs(mu - ld) < 1e-20) { num = mu * time; den = 1.0 + num; }
    else if (a < 0.) { const double A = (ld > 0.) ? std::exp(expo + std::log(ld)) : 0.0; num = ld - A; den = mu - A; }
    else {
      const double Al = (ld > 0.) ? std::exp(expo + std::log(ld)) : 0.0, Am = (mu > 0.) ? std::exp(expo + std::log(mu)) : 0.0;
      num = ld - Al; den = ld - Am;
    }
    beta = (std::fabs(den) > 0.) ? num / den : 1e-8;
    if (std::isnan(beta)) return RSG_FAIL;
    if (beta > 1.) { if (beta < 1.0 + 1e-3) beta = 1.0; else return RSG_FAIL; }
    if (beta < 0.) { if (beta > -1e-4) beta = 0.0; else return RSG_FAIL; }
  }
  const double move = 1.0 - (double)L / ((double)L + 1.0);
  const double rM = R->rM, rD = R->rD, rI = R->rI;
  // e1_model_transitions_AF in e2_GLOBAL mode (src/e1_model.c:784-868)
  t[0] = (1.0 - beta) * (1.0 - gM);              t[1] = (1.0 - beta) * gM;
  t[2] = (beta == 1.0) ? 1.0 - move : beta;
  t[4] = (1.0 - rM) * (1.0 - beta) * (1.0 - gM) + rM;  t[5] = (1.0 - rM) * (1.0 - beta) * gM;
  t[6] = ((1.0 - rM) * beta == 1.0) ? 1.0 - move : (1.0 - rM) * beta;
  t[8] = (1.0 - rD) * (1.0 - beta) * (1.0 - gD);  t[9] = (1.0 - rD) * (1.0 - beta) * gD + rD;
  t[10] = ((1.0 - rD) * beta == 1.0) ? 1.0 - move : (1.0 - rD) * beta;
  t[12] = (1.0 - rI) * (1.0 - beta) * (1.0 - gI);  t[13] = (1.0 - rI) * (1.0 - beta) * gI;
  t[14] = ((1.0 - rI) * beta == 1.0) ? 1.0 - move : (1.0 - rI) * beta + rI;
  for (int r = 0; r < 4; r++) {                     // e1_model_RenormStateE (:214-270)
    double s = t[4 * r] + t[4 * r + 1] + t[4 * r + 2];
    s = (s > 0.) ? 1.0 / s : 1.0;
    for (int k = 0; k < 3; k++) t[4 * r + k] *= s;
    t[4 * r + 3] = 0.0;
  }
  return RSG_OK;
}

extern "C" int rsg_cov_GenerateAlignment(rsg_ctx *c, const rsg_null_job *job, const uint8_t *root, const int *ct,
                                         const double *model16, const rsg_e1_rates *rates, const double *insf,
                                         uint8_t **ret_leaves, int64_t *ret_alen, int **ret_ct, char *errbuf)
{
  if (!c || !job || !ret_leaves || !ret_alen) return rsg_fail(errbuf, "rsg_cov_GenerateAlignment: bad arguments");
  *ret_leaves = nullptr; *ret_alen = 0;
  if (ret_ct) *ret_ct = nullptr;
  const int64_t N = job->nleaf, L0 = job->alen;
  if (N < 2 || L0 < 1) return rsg_fail(errbuf, "simulation needs >= 2 leaves and >= 1 column");
  if (root)
    for (int64_t k = 0; k < L0; k++)
      if (root[k] > 3) return rsg_fail(errbuf, "rsg_cov_GenerateAlignment: root residue %lld is not canonical", (long long)k);
  std::vector<uint8_t> ung((size_t)(N * L0));
  const bool pairs = ct && model16;
  if (pairs) { if (rsg_sim_pairs(c, job, ct, model16, root, ung.data(), errbuf) != RSG_OK) return RSG_FAIL; }
  else if (rsg_sim_one(c, job, root, ung.data(), errbuf) != RSG_OK) return RSG_FAIL;

  auto finish = [&](const uint8_t *leaves, int64_t F, const std::vector<int> &ctf) -> int {
    uint8_t *o = (uint8_t *)malloc((size_t)(N * F));
    if (!o) return rsg_fail(errbuf, "out of memory");
    memcpy(o, leaves, (size_t)(N * F));
    if (ret_ct) {
      int *q = (int *)malloc(sizeof(int) * (size_t)F);
      if (!q) { free(o); return rsg_fail(errbuf, "out of memory"); }
      memcpy(q, ctf.data(), sizeof(int) * (size_t)F);
      *ret_ct = q;
    }
    *ret_leaves = o; *ret_alen = F;
    return RSG_OK;
  };
  if (!rates) {                                      // noindels
    std::vector<int> ctf((size_t)L0, -1);
    if (ct) for (int64_t k = 0; k < L0; k++) ctf[(size_t)k] = ct[k];
    return finish(ung.data(), L0, ctf);
  }
  if (!(rates->rI >= 0. && rates->rI < 1. && rates->rM >= 0. && rates->rM < 1. && rates->rD >= 0. && rates->rD < 1.))
    return rsg_fail(errbuf, "rsg_cov_GenerateAlignment: rI, rM, rD must lie in [0, 1)");

  // branches in tree order: b = 2 v + side
  const int nb = (int)(2 * (N - 1));
  std::vector<uint32_t> th((size_t)nb * 8);
  std::vector<int> rows((size_t)nb), child((size_t)nb);
  double expect = 0.;                                 // events per column, largest over the branches
  for (int64_t v = 0; v < N - 1; v++)
    for (int side = 0; side < 2; side++) {
      const int b = (int)(2 * v + side), d = side ? job->right[v] : job->left[v];
      const double time = side ? job->rd[v] : job->ld[v];
      double t[16];
      if (rsg_e1_transitions(rates, time, L0, t) != RSG_OK)
        return rsg_fail(errbuf, "rsg_cov_GenerateAlignment: e1 transitions failed for branch %d (time %g)", b, time);
      for (int r = 0; r < 4; r++) {
        const double s = t[4 * r] + t[4 * r + 1] + t[4 * r + 2];
        th[(size_t)b * 8 + 2 * r] = th32(t[4 * r] / s);
        th[(size_t)b * 8 + 2 * r + 1] = th32((t[4 * r] + t[4 * r + 1]) / s);
      }
      rows[(size_t)b] = (d > 0) ? d : (int)(N - 1 - d);
      child[(size_t)b] = d;
      const double ext = t[14] < 1. ? t[2] / (1. - t[14]) : 1e9;
      expect = std::max(expect, t[1] + ext);
    }
  uint32_t ith[3];
  {
    const double *f = insf ? insf : job->model + 36;     // insertion residues: the given frequencies or pi
    double cum = 0., tot = f[0] + f[1] + f[2] + f[3];
    if (!(tot > 0.)) return rsg_fail(errbuf, "rsg_cov_GenerateAlignment: insertion frequencies must be positive");
    for (int k = 0; k < 3; k++) { cum += f[k]; ith[k] = th32(cum / tot); }
  }

  HIPCHK(hipSetDevice(c->device));
  int64_t Lmax = 2 * L0 + 64;
  int cap = 64;
  std::vector<int> cnt((size_t)nb), ev;
  std::vector<int64_t> Lb((size_t)nb);
  int64_t F = 0;
  for (int attempt = 0;; attempt++) {
    if (attempt > 24) return rsg_fail(errbuf, "rsg_cov_GenerateAlignment: indel walks do not fit (rates too high)");
    cap = std::max(cap, (int)std::min<double>(1 << 26, 64. + 4. * expect * (double)(Lmax + 1)));
    DevMem dm;
    uint32_t *d_th; int *d_rows, *d_ev, *d_cnt;
    if (!dm.get(&d_th, th.size()) || !dm.get(&d_rows, rows.size()) || !dm.get(&d_ev, (size_t)nb * cap) ||
        !dm.get(&d_cnt, (size_t)nb))
      return rsg_fail(errbuf, "rsg_cov_GenerateAlignment: out of device memory");
    HIPCHK(hipMemcpyAsync(d_th, th.data(), sizeof(uint32_t) * th.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d_rows, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice, c->stream));
    int tm; rsg_tstart(c, "indel_walk", &tm);
    hipLaunchKernelGGL(k_indel_walk, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, c->stream, nb, d_th, d_rows, job->seed,
                       job->stream, Lmax, cap, d_ev, d_cnt);
    rsg_tstop(c, tm);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(cnt.data(), d_cnt, sizeof(int) * cnt.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    bool over = false;
    for (int v : cnt) if (v > cap) { over = true; break; }
    if (over) { cap *= 2; continue; }
    ev.resize((size_t)nb * cap);
    HIPCHK(hipMemcpy(ev.data(), d_ev, sizeof(int) * ev.size(), hipMemcpyDeviceToHost));
    // the scan over branches in tree order: each walks the columns present when it starts
    int64_t len = L0;
    bool fits = true;
    for (int b = 0; b < nb && fits; b++) {
      Lb[(size_t)b] = len;
      if (len > Lmax) { fits = false; break; }
      int64_t ins = 0;
      for (int k = 0; k < cnt[(size_t)b]; k++) {
        const int e = ev[(size_t)b * cap + k];
        if ((e >> 1) > len) break;                     // past this branch's end
        ins += e & 1;
      }
      len += ins;
    }
    if (!fits) { Lmax = 2 * Lmax + 64; continue; }
    F = len;
    break;
  }
  if (F > INT32_MAX || N * F > ((int64_t)1 << 40)) return rsg_fail(errbuf, "rsg_cov_GenerateAlignment: alignment too large");

  // chronological events: insertions at index q of the alignment of that moment, deletions of
  // the cell at index x; leaf branches' events also go to the leaves
  struct Ev { int b, kind; int64_t q; int ord; };
  std::vector<Ev> evs;
  for (int b = 0; b < nb; b++) {
    int nins = 0;
    for (int k = 0; k < cnt[(size_t)b]; k++) {
      const int e = ev[(size_t)b * cap + k];
      const int64_t i = e >> 1;
      if (i > Lb[(size_t)b]) break;
      if (e & 1) { evs.push_back(Ev{b, 1, i + nins, nins}); nins++; }
      else evs.push_back(Ev{b, 0, i + nins - 1, -1});
    }
  }
  // final column of every insertion: in reverse, the (q + 1)-th slot still free
  std::vector<int> fpos(evs.size(), -1);
  Fenwick fr((int)F);
  for (int64_t k = 0; k < F; k++) fr.add((int)k, 1);
  for (int64_t k = (int64_t)evs.size() - 1; k >= 0; k--) {
    if (evs[(size_t)k].kind != 1) continue;
    const int p = fr.kth((int)evs[(size_t)k].q + 1);
    fpos[(size_t)k] = p;
    fr.add(p, -1);
  }
  std::vector<int> src((size_t)F, -1), orig((size_t)L0);
  {
    int64_t j = 0;
    for (int64_t k = 0; k < F; k++)
      if (fr.kth((int)(j + 1)) == k && j < L0) { orig[(size_t)j] = (int)k; src[(size_t)k] = (int)j; j++; }
    if (j != L0) return rsg_fail(errbuf, "rsg_cov_GenerateAlignment: column composition failed");
  }
  // deletions: the (x + 1)-th column present at that moment
  std::vector<int4> sc;
  Fenwick pr((int)F);
  for (int64_t j = 0; j < L0; j++) pr.add(orig[(size_t)j], 1);
  for (size_t k = 0; k < evs.size(); k++) {
    const Ev &e = evs[k];
    const int d = child[(size_t)e.b];
    if (e.kind == 1) {
      pr.add(fpos[k], 1);
      if (d <= 0) sc.push_back(make_int4(-d, fpos[k], rows[(size_t)e.b], e.ord));
    } else if (d <= 0) {
      sc.push_back(make_int4(-d, pr.kth((int)e.q + 1), rows[(size_t)e.b], -1));
    }
  }
  std::vector<int> ctf((size_t)F, -1);
  if (ct) for (int64_t j = 0; j < L0; j++) if (ct[j] >= 0) ctf[(size_t)orig[(size_t)j]] = orig[(size_t)ct[j]];

  DevMem dm;
  uint8_t *d_ung, *d_out; int *d_src; int4 *d_sc; uint32_t *d_ith;
  if (!dm.get(&d_ung, (size_t)(N * L0)) || !dm.get(&d_out, (size_t)(N * F)) || !dm.get(&d_src, (size_t)F) ||
      !dm.get(&d_sc, sc.size()) || !dm.get(&d_ith, 3))
    return rsg_fail(errbuf, "rsg_cov_GenerateAlignment: out of device memory");
  HIPCHK(hipMemcpyAsync(d_ung, ung.data(), (size_t)(N * L0), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_src, src.data(), sizeof(int) * (size_t)F, hipMemcpyHostToDevice, c->stream));
  if (!sc.empty()) HIPCHK(hipMemcpyAsync(d_sc, sc.data(), sizeof(int4) * sc.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_ith, ith, sizeof(ith), hipMemcpyHostToDevice, c->stream));
  int tm; rsg_tstart(c, "indel_build", &tm);
  hipLaunchKernelGGL(k_indel_gather, dim3((unsigned)((N * F + 255) / 256)), dim3(256), 0, c->stream, d_ung, N, L0, d_src, F, d_out);
  if (!sc.empty())
    hipLaunchKernelGGL(k_indel_scatter, dim3((unsigned)((sc.size() + 255) / 256)), dim3(256), 0, c->stream, d_sc, (int)sc.size(),
                       F, job->seed, job->stream, d_ith, d_out);
  rsg_tstop(c, tm);
  HIPCHK(hipGetLastError());
  std::vector<uint8_t> out((size_t)(N * F));
  HIPCHK(hipMemcpyAsync(out.data(), d_out, out.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return finish(out.data(), F, ctf);
}
