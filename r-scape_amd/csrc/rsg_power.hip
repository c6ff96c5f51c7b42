// rsg_power.hip -- substitutions on the tree, the power of base pairs to covary, the hit list
// and the output files of a significance run.
//
// Reference (R-scape v2.6.11):
//   Tree_Substitutions (src/msatree.c:1423-1555): Fitch ancestral states (the same passes as
//     R-scape's null, rsg_fitch.hip), then per column the branches that substitute (nsubs) and
//     per column pair the branches that substitute at both columns (ndouble) or at either
//     (njoin).  Here a branch x column code is formed on the device -- 0 a substitution, 1 no
//     substitution, 4 neither (a gap or other code at an end, without --powergaps) -- and the
//     pair counts are the int8 one-hot Gram of that branch alignment: ndouble = c[0][0],
//     njoin = c[0][0] + c[0][1] + c[1][0] (counts of the 4 x 4 table of codes 0 and 1).
//   power_Read (src/power.c:532-650), power_SPAIR_Create + power_SPAIR_AddMI (:65-217),
//   cov_CreateHitList's selection (src/covariation.c:778-1006), cov_WriteHitList /
//   cov_WriteRankedHitList (:1139-1406), power_SPAIR_Write (src/power.c:261-378) and the
//   survival plot data of cov_WriteHistogram (src/covariation.c:1637-1673, 1870-1890).
#include "rsg_internal.h"
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cctype>
#include <vector>

namespace {

__device__ __forceinline__ int64_t row_of_child(int d, int64_t N) { return d > 0 ? N + d : -d; }

// branch b = 2 v + side, column c: 0 = substitution, 1 = none, 4 = not counted
__global__ void k_branch_codes(int64_t N, int64_t L, const int *__restrict__ left, const int *__restrict__ right,
                               const uint8_t *__restrict__ all, int includegaps, uint8_t *__restrict__ codes)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * (N - 1) * L) return;
  const int64_t b = e / L, c = e - b * L;
  const int v = (int)(b >> 1);
  const int d = (b & 1) ? right[v] : left[v];
  const unsigned pa = all[(N + v) * L + c], ch = all[row_of_child(d, N) * L + c];
  uint8_t code;
  if (includegaps) code = (pa != ch) ? 0 : 1;
  else             code = (pa < 4u && ch < 4u) ? ((pa != ch) ? 0 : 1) : 4;
  codes[e] = code;
}

// nsubs[c]: branches with a substitution at column c (one thread per column)
__global__ void k_nsubs(int64_t nb, int64_t L, const uint8_t *__restrict__ codes, int32_t *__restrict__ nsubs)
{
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= L) return;
  int32_t n = 0;
  for (int64_t b = 0; b < nb; b++) n += (codes[b * L + c] == 0);
  nsubs[c] = n;
}

// ndouble / njoin of every pair i < j at i * L + j from the code Gram's records
template <int KIND>
__global__ void k_join_double(const void *__restrict__ cnt, int64_t L, int32_t *__restrict__ ndouble, int32_t *__restrict__ njoin)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L * L) return;
  const int64_t i = e / L, j = e - i * L;
  int32_t dbl = 0, jn = 0;
  if (i < j) {
    const int64_t r = rsg_rec(i, j, L) * 16;
    uint32_t c00, c01, c10;
    if (KIND == CNT_U16) { const uint16_t *p = (const uint16_t *)cnt + r; c00 = p[0]; c01 = p[1]; c10 = p[4]; }
    else                 { const uint32_t *p = (const uint32_t *)cnt + r; c00 = p[0]; c01 = p[1]; c10 = p[4]; }
    dbl = (int32_t)c00; jn = (int32_t)(c00 + c01 + c10);
  }
  if (ndouble) ndouble[e] = dbl;
  if (njoin) njoin[e] = jn;
}

// per pair i < j, in the reference's pair order n (i ascending, then j): nseff, ngap and the
// Watson-Crick and G:U fractions of pp (power_SPAIR_AddMI: pp = (c + 1e-10) / (nseff + 16e-10))
template <int KIND>
__global__ void k_pair_summary(const void *__restrict__ cnt, int64_t L, double wsum, double4 *__restrict__ out)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L * L) return;
  const int64_t i = e / L, j = e - i * L;
  if (i >= j) return;
  double c[16];
  const int64_t r = rsg_rec(i, j, L) * 16;
  for (int k = 0; k < 16; k++)
    c[k] = (KIND == CNT_U16) ? (double)((const uint16_t *)cnt)[r + k]
         : (KIND == CNT_U32) ? (double)((const uint32_t *)cnt)[r + k] : (double)((const float *)cnt)[r + k];
  double ns = 0.;
  for (int k = 0; k < 16; k++) ns += c[k];
  const double den = ns + 16e-10;
  const double wcf = ((c[0 * 4 + 3] + 1e-10) + (c[3 * 4 + 0] + 1e-10) + (c[1 * 4 + 2] + 1e-10) + (c[2 * 4 + 1] + 1e-10)) / den;
  const double wob = ((c[2 * 4 + 3] + 1e-10) + (c[3 * 4 + 2] + 1e-10)) / den;
  const int64_t n = i * L - i * (i + 1) / 2 + (j - i - 1);
  out[n] = make_double4(ns, wsum - ns, wcf, wob);
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

extern "C" int rsg_tree_Substitutions(rsg_ctx *c, const uint8_t *ax, int64_t N, int64_t L, const int *left, const int *right,
                                      uint64_t seed, int includegaps, int32_t *nsubs, int32_t *ndouble, int32_t *njoin,
                                      char *errbuf)
{
  if (!c || !ax || !left || !right) return rsg_fail(errbuf, "rsg_tree_Substitutions: bad arguments");
  HIPCHK(hipSetDevice(c->device));
  FitchPlan P;
  if (rsg_fitch_plan(N, L, left, right, P, errbuf) != RSG_OK) return RSG_FAIL;
  const int64_t nrow = 2 * N - 1, m = P.m, nb = 2 * m;
  Buf B;
  uint8_t *d_ax, *d_mask, *d_all, *d_codes; int *d_left, *d_right, *d_parent, *d_nodes; uint64_t *d_seed; int32_t *d_ns;
  if (!B.get(&d_ax, (size_t)(N * L)) || !B.get(&d_mask, (size_t)(nrow * L)) || !B.get(&d_all, (size_t)(nrow * L)) ||
      !B.get(&d_codes, (size_t)(nb * L)) || !B.get(&d_left, (size_t)m) || !B.get(&d_right, (size_t)m) ||
      !B.get(&d_parent, (size_t)m) || !B.get(&d_nodes, (size_t)m) || !B.get(&d_seed, 1) || !B.get(&d_ns, (size_t)L))
    return rsg_fail(errbuf, "rsg_tree_Substitutions: out of device memory");
  HIPCHK(hipMemcpyAsync(d_ax, ax, (size_t)(N * L), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_left, left, sizeof(int) * m, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_right, right, sizeof(int) * m, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_parent, P.parent.data(), sizeof(int) * m, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_nodes, P.nodes.data(), sizeof(int) * m, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_seed, &seed, sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
  FitchDev D{d_ax, d_left, d_right, d_parent, d_nodes, nullptr, d_seed, d_mask, d_all, nullptr, nullptr, nullptr, nullptr};
  int t; rsg_tstart(c, "substitutions", &t);
  if (rsg_fitch_ancestral(c, P, D, 1) != RSG_OK) return rsg_fail(errbuf, "rsg_tree_Substitutions: Fitch launch failed");
  hipLaunchKernelGGL(k_branch_codes, dim3((unsigned)((nb * L + 255) / 256)), dim3(256), 0, c->stream, N, L, d_left, d_right,
                     d_all, includegaps, d_codes);
  hipLaunchKernelGGL(k_nsubs, dim3((unsigned)((L + 255) / 256)), dim3(256), 0, c->stream, nb, L, d_codes, d_ns);
  HIPCHK(hipGetLastError());
  int st = RSG_OK;
  if ((ndouble || njoin) && L >= 2) {
    // the branch alignment's one-hot Gram (codes 0 and 1 are the counted symbols)
    rsg_msa bm;
    bm.N = nb; bm.L = L; bm.Npad = rsg_npad(nb); bm.Lpad = ((L + 63) / 64) * 64;
    bm.d_ax = d_codes; bm.wsum = (double)nb;
    const int kind = rsg_int_kind(nb);
    int32_t *d_dbl = nullptr, *d_jn = nullptr;
    if (!B.get(&bm.d_axT, (size_t)(bm.Lpad * bm.Npad)) || !B.get(&bm.d_xt, (size_t)(bm.Lpad * bm.Npad)) ||
        !B.get((uint8_t **)&bm.d_cnt, (size_t)rsg_rec_bytes(kind) * (size_t)(L * L)) ||
        !B.get(&d_dbl, (size_t)(L * L)) || !B.get(&d_jn, (size_t)(L * L)))
      return rsg_fail(errbuf, "rsg_tree_Substitutions: out of device memory");
    bm.cnt_kind = kind;
    if (rsg_pack(c, &bm, errbuf) != RSG_OK || rsg_gram_counts_i8(c, &bm, bm.d_cnt, kind, nullptr, nullptr, errbuf) != RSG_OK)
      st = RSG_FAIL;
    if (st == RSG_OK) {
      const dim3 g((unsigned)((L * L + 255) / 256));
      if (kind == CNT_U16) hipLaunchKernelGGL(k_join_double<CNT_U16>, g, dim3(256), 0, c->stream, bm.d_cnt, L, d_dbl, d_jn);
      else                 hipLaunchKernelGGL(k_join_double<CNT_U32>, g, dim3(256), 0, c->stream, bm.d_cnt, L, d_dbl, d_jn);
      HIPCHK(hipGetLastError());
      if (ndouble) HIPCHK(hipMemcpyAsync(ndouble, d_dbl, sizeof(int32_t) * (size_t)(L * L), hipMemcpyDeviceToHost, c->stream));
      if (njoin) HIPCHK(hipMemcpyAsync(njoin, d_jn, sizeof(int32_t) * (size_t)(L * L), hipMemcpyDeviceToHost, c->stream));
    }
  } else {
    if (ndouble) memset(ndouble, 0, sizeof(int32_t) * (size_t)(L * L));
    if (njoin) memset(njoin, 0, sizeof(int32_t) * (size_t)(L * L));
  }
  rsg_tstop(c, t);
  if (st != RSG_OK) return RSG_FAIL;
  if (nsubs) HIPCHK(hipMemcpyAsync(nsubs, d_ns, sizeof(int32_t) * (size_t)L, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return RSG_OK;
}

// ---------------------------------------------------------------------------------- power
extern "C" int rsg_power_Read(const char *path, rsg_power **ret, char *errbuf)
{
  *ret = nullptr;
  FILE *fp = fopen(path, "r");
  if (!fp) return rsg_fail(errbuf, "open failed: %s", path);
  std::vector<double> subs, prob;
  bool reached = false;
  char line[4096];
  while (fgets(line, sizeof line, fp)) {
    // fields split at whitespace: the first is the substitution count, the second the power
    char f[2][4096] = {{0}, {0}};
    int idx = 0, len = 0;
    for (const char *p = line; *p && *p != '\n' && idx < 2; p++) {
      if (!isspace((unsigned char)*p)) f[idx][len++] = *p;
      else { f[idx][len] = 0; len = 0; idx++; }
    }
    if (idx < 2) f[idx][len] = 0;
    double s = atof(f[0]), pr = atof(f[1]);
    if (pr < 0.) pr = 0.;
    if (pr > 1.) { pr = 1.; reached = true; }
    if (reached) pr = 1.;
    subs.push_back(s); prob.push_back(pr);
  }
  fclose(fp);
  rsg_power *p = (rsg_power *)calloc(1, sizeof(rsg_power));
  if (!p) return rsg_fail(errbuf, "out of memory");
  p->ns = (int)subs.size();
  p->subs = (double *)malloc(sizeof(double) * std::max<size_t>(subs.size(), 1));
  p->prob = (double *)malloc(sizeof(double) * std::max<size_t>(prob.size(), 1));
  if (!p->subs || !p->prob) { rsg_power_Destroy(p); return rsg_fail(errbuf, "out of memory"); }
  std::copy(subs.begin(), subs.end(), p->subs);
  std::copy(prob.begin(), prob.end(), p->prob);
  *ret = p;
  return RSG_OK;
}

extern "C" void rsg_power_Destroy(rsg_power *p)
{
  if (!p) return;
  free(p->subs); free(p->prob); free(p);
}

// the probability of the last table row whose substitution count the pair exceeds
extern "C" double rsg_power_Prob(const rsg_power *p, int64_t nsubs)
{
  double prob = 0.;
  if (!p) return 0.;
  for (int s = 0; s < p->ns; s++) {
    if ((double)nsubs > p->subs[s]) prob = p->prob[s];
    else break;
  }
  return prob;
}

extern "C" int rsg_power_SPAIR_Create(rsg_ctx *c, rsg_msa *m, const int *msamap, const rsg_power *power, int powertype,
                                      const int *ct, const int32_t *nsubs, const int32_t *njoin, const int32_t *ndouble,
                                      rsg_spair *spair, char *errbuf)
{
  if (!c || !m || !spair) return rsg_fail(errbuf, "rsg_power_SPAIR_Create: bad arguments");
  if (!m->have_probs || !m->d_cnt) return rsg_fail(errbuf, "rsg_power_SPAIR_Create: corr_Probs has not run");
  if (!nsubs && !njoin && !ndouble) return rsg_fail(errbuf, "need some kind of substitutions!");
  HIPCHK(hipSetDevice(c->device));
  const int64_t L = m->L, dim = L * (L - 1) / 2;
  std::vector<double4> sum((size_t)std::max<int64_t>(dim, 1));
  if (dim > 0) {
    double4 *d_sum = nullptr;
    HIPCHK(hipMalloc(&d_sum, sizeof(double4) * (size_t)dim));
    const dim3 g((unsigned)((L * L + 255) / 256));
    if (m->cnt_kind == CNT_U16)      hipLaunchKernelGGL(k_pair_summary<CNT_U16>, g, dim3(256), 0, c->stream, m->d_cnt, L, rsg_count_wsum(m), d_sum);
    else if (m->cnt_kind == CNT_U32) hipLaunchKernelGGL(k_pair_summary<CNT_U32>, g, dim3(256), 0, c->stream, m->d_cnt, L, rsg_count_wsum(m), d_sum);
    else                             hipLaunchKernelGGL(k_pair_summary<CNT_F32>, g, dim3(256), 0, c->stream, m->d_cnt, L, rsg_count_wsum(m), d_sum);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(sum.data(), d_sum, sizeof(double4) * (size_t)dim, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(d_sum);
    if (e != hipSuccess) return rsg_fail(errbuf, "rsg_power_SPAIR_Create: %s", hipGetErrorString(e));
  }
  int64_t n = 0;
  for (int64_t i = 0; i < L - 1; i++)
    for (int64_t j = i + 1; j < L; j++, n++) {
      rsg_spair &s = spair[n];
      s.i = i; s.j = j;
      s.iabs = (msamap ? msamap[i] : i) + 1; s.jabs = (msamap ? msamap[j] : j) + 1;
      s.nseff = sum[(size_t)n].x; s.ngap = sum[(size_t)n].y;
      s.wcf_f = sum[(size_t)n].z; s.wobble_f = sum[(size_t)n].w;
      s.noncan_f = 1.0 - s.wcf_f - s.wobble_f;
      s.powertype = ndouble ? RSG_DOUBLE_SUBS : (njoin ? RSG_JOIN_SUBS : RSG_SINGLE_SUBS);
      s.nsubs = nsubs ? (int64_t)nsubs[i] + nsubs[j] : 0;
      s.nsubs_join = njoin ? njoin[i * L + j] : 0;
      s.nsubs_double = ndouble ? ndouble[i * L + j] : 0;
      s.power = 0.; s.power_join = -1.; s.power_double = -1.;
      if (power) {
        if (nsubs)   s.power = rsg_power_Prob(power, s.nsubs);
        if (njoin)   s.power_join = rsg_power_Prob(power, s.nsubs_join);
        if (ndouble) s.power_double = rsg_power_Prob(power, s.nsubs_double);
      }
      s.bptype_given = (ct && ct[i] == (int)j && ct[j] == (int)i) ? RSG_BP_WWC : RSG_BP_NONE;
      s.covary = 0;
      s.sc = -1.; s.Pval = -1.; s.Eval = -1.;
    }
  (void)powertype;
  return RSG_OK;
}

// ---------------------------------------------------------------------------------- hit list
extern "C" int rsg_cov_CreateHitList(rsg_ctx *c, rsg_msa *m, const rsg_hist *null, const double *survfit, int samplesize,
                                     const int *ct, const uint8_t *bptype, int64_t Nt, int64_t Nb, double expBP,
                                     double thresh, rsg_spair *spair, rsg_hit **ret_hits, int *ret_nhit, char *errbuf)
{
  *ret_hits = nullptr; *ret_nhit = 0;
  if (!c || !m || !ret_hits) return rsg_fail(errbuf, "rsg_cov_CreateHitList: bad arguments");
  if (!m->have_cov) return rsg_fail(errbuf, "no covariation matrix");
  const int64_t L = m->L;
  std::vector<double> cov((size_t)(L * L)), pv;
  if (rsg_corr_GetCOV(c, m, cov.data(), nullptr, nullptr, errbuf) != RSG_OK) return RSG_FAIL;
  if (null) {                                   // p-values: cov2evalue(cov, 1, ...) for every pair
    pv.resize((size_t)(L * L));
    if (rsg_cov_Evalues(c, m, null, survfit, 1, pv.data(), errbuf) != RSG_OK) return RSG_FAIL;
  }
  std::vector<rsg_hit> hits;
  int64_t n = 0;
  int h = 0;
  for (int64_t i = 0; i < L - 1; i++)
    for (int64_t j = i + 1; j < L; j++, n++) {
      const int is_compatible = (ct && ct[i] >= 0) ? 0 : 1;     // the reference tests position i twice
      int bpt = bptype ? bptype[i * L + j] : ((ct && ct[i] == (int)j && ct[j] == (int)i) ? RSG_BP_WWC : RSG_BP_NONE);
      int isbp = 0;
      switch (samplesize) {
      case RSG_SAMPLE_CONTACTS: isbp = bpt < RSG_BP_NONE; break;
      case RSG_SAMPLE_BP:       isbp = bpt < RSG_BP_STACKED; break;
      case RSG_SAMPLE_WC:       isbp = bpt == RSG_BP_WWC; break;
      default:                  isbp = 0; break;
      }
      const double sc = cov[(size_t)(i * L + j)];
      double pval = 0., eval = 0.;
      if (null) {
        pval = pv[(size_t)(i * L + j)];
        if (isbp)                  eval = pval * (double)Nb;
        else if ((double)h < expBP) eval = pval * expBP;
        else                       eval = pval * (double)Nt;
      }
      if (eval < thresh || thresh > 1000.) {
        rsg_hit H;
        H.i = i; H.j = j; H.sc = sc; H.Eval = eval; H.pval = pval;
        H.nsubs = 0; H.power = 0.;
        if (spair) {
          switch (spair[n].powertype) {
          case RSG_JOIN_SUBS:   H.nsubs = spair[n].nsubs_join;   H.power = spair[n].power_join;   break;
          case RSG_DOUBLE_SUBS: H.nsubs = spair[n].nsubs_double; H.power = spair[n].power_double; break;
          default:              H.nsubs = spair[n].nsubs;        H.power = spair[n].power;        break;
          }
          spair[n].covary = 1;
        }
        H.bptype = bpt; H.is_compatible = is_compatible;
        hits.push_back(H);
        h++;
      }
      if (spair) { spair[n].sc = sc; spair[n].Eval = eval; spair[n].Pval = pval; }
    }
  rsg_hit *out = (rsg_hit *)malloc(sizeof(rsg_hit) * std::max<size_t>(hits.size(), 1));
  if (!out) return rsg_fail(errbuf, "out of memory");
  std::copy(hits.begin(), hits.end(), out);
  *ret_hits = out; *ret_nhit = (int)hits.size();
  return RSG_OK;
}

// ---------------------------------------------------------------------------------- files
namespace {
void pair_canonicals(const rsg_spair *spair, int64_t alen, int64_t i, int64_t j, double *can, double *nseff, double *nseff_f)
{
  if (!spair) { *can = 0.; *nseff = 0.; *nseff_f = 0.; return; }
  const int64_t n = i * alen - i * (i + 1) / 2 + (j - i - 1);
  const rsg_spair &s = spair[n];
  *can = 100. * (s.wcf_f + s.wobble_f);
  *nseff = s.nseff;
  *nseff_f = (s.nseff + s.ngap) ? 100. * (s.nseff / (s.nseff + s.ngap)) : 0.;
}

const char *hit_mark(const rsg_hit &h)
{
  if (h.bptype == RSG_BP_WWC) return "*";
  if (h.bptype < RSG_BP_STACKED) return "**";
  if (h.bptype < RSG_BP_NONE && h.is_compatible) return "c~";
  if (h.bptype < RSG_BP_NONE) return "c";
  if (h.is_compatible) return "~";
  return " ";
}

// hit_sorted_by_eval / hit_sorted_by_score (src/covariation.c:1319-1352): E-value ascending
// (score descending); on a tie the reference's comparator returns dir2 when the i and j
// directions differ, else dir1, with dir1 = (a.i < b.i ? 1 : -1), dir2 = (a.j < b.j ? 1 : -1)
int cmp_ref(const rsg_hit *a, const rsg_hit *b, bool by_score)
{
  if (by_score) { if (a->sc < b->sc) return 1; if (a->sc > b->sc) return -1; }
  else          { if (a->Eval > b->Eval) return 1; if (a->Eval < b->Eval) return -1; }
  const int dir1 = a->i < b->i ? 1 : -1, dir2 = a->j < b->j ? 1 : -1;
  return (dir1 != dir2) ? dir2 : dir1;
}

// The reference sorts with qsort; the comparator is not a consistent order on ties, so the
// result depends on the algorithm.  This is the top-down merge sort of GNU libc's qsort (halves
// n / 2 and n - n / 2, the left element taken while cmp <= 0).
void merge_sort(std::vector<const rsg_hit *> &v, std::vector<const rsg_hit *> &tmp, size_t lo, size_t n, bool by_score)
{
  if (n <= 1) return;
  const size_t n1 = n / 2, n2 = n - n1;
  merge_sort(v, tmp, lo, n1, by_score);
  merge_sort(v, tmp, lo + n1, n2, by_score);
  size_t a = lo, b = lo + n1, k = 0;
  while (a < lo + n1 && b < lo + n)
    tmp[k++] = (cmp_ref(v[a], v[b], by_score) <= 0) ? v[a++] : v[b++];
  while (a < lo + n1) tmp[k++] = v[a++];
  while (b < lo + n) tmp[k++] = v[b++];
  for (size_t q = 0; q < n; q++) v[lo + q] = tmp[q];
}
} // namespace

extern "C" int rsg_cov_WriteHitList(const char *path, int ranked, int by_score, const char *covtype, double thresh,
                                    double cov_min, double cov_max, int tnum, int64_t Nb, int64_t Nt, const rsg_hit *hits,
                                    int nhit, const rsg_spair *spair, int64_t alen, const int *msamap, int firstpos,
                                    int samplesize, char *errbuf)
{
  FILE *fp = fopen(path, "w");
  if (!fp) return rsg_fail(errbuf, "could not open %s", path);
  // the summary line of cov_CreateHitList (found / true positives)
  int f = 0, tf = 0;
  for (int h = 0; h < nhit; h++) {
    f++;
    int isbp;
    switch (samplesize) {
    case RSG_SAMPLE_BP: isbp = hits[h].bptype < RSG_BP_STACKED; break;
    case RSG_SAMPLE_WC: isbp = hits[h].bptype == RSG_BP_WWC; break;
    default:            isbp = hits[h].bptype < RSG_BP_NONE; break;      // contacts, and all
    }
    if (isbp) tf++;
  }
  const int fpn = f - tf;
  const double sen = (tnum > 0) ? 100. * (double)tf / (double)tnum : 0.0;
  const double ppv = (f > 0) ? 100. * (double)tf / (double)f : 0.0;
  const double F = (sen + ppv > 0.) ? 2.0 * sen * ppv / (sen + ppv) : 0.0;
  fprintf(fp, "#\n# Method Target_E-val [cov_min,cov_max] [FP | TP True Found | Sen PPV F] \n");
  auto pos = [&](int64_t k) { return (msamap ? msamap[k] : (int)k) + firstpos; };
  if (!ranked) {
    fprintf(fp, "# %s    %g           [%.2f,%.2f]    [%d | %d %d %d | %.2f %.2f %.2f] \n#\n", covtype, thresh, cov_min, cov_max,
            fpn, tf, tnum, f, sen, ppv, F);
    if (nhit == 0) {
      fprintf(fp, "#-------------------------------------------------------------------------------------------------------\n");
      fprintf(fp, "no significant pairs\n");
    } else {
      fprintf(fp, "# #tests = %lld (base pairs) %lld (others)\n", (long long)Nb, (long long)Nt);
      fprintf(fp, "# in_given  left_pos       right_pos        score      E-value          pvalue      substitutions       power   nseff(%%)       %% canonical_pairs\n");
      fprintf(fp, "#-------------------------------------------------------------------------------------------------------------------------------------------------\n");
    }
    for (int h = 0; h < nhit; h++) {
      double can, ns, nsf;
      pair_canonicals(spair, alen, hits[h].i, hits[h].j, &can, &ns, &nsf);
      fprintf(fp, "%s\t%10d\t%10d\t%.5f\t%g\t%g\t%lld\t\t%.2f\t%.2f (%.1f%%)\t%.1f\n", hit_mark(hits[h]), pos(hits[h].i),
              pos(hits[h].j), hits[h].sc, hits[h].Eval, hits[h].pval, (long long)hits[h].nsubs, hits[h].power, ns, nsf, can);
    }
  } else {
    fprintf(fp, "# %s    %g         [%.2f,%.2f]     [%d | %d %d %d | %.2f %.2f %.2f] \n#\n", covtype, thresh, cov_min, cov_max,
            fpn, tf, tnum, f, sen, ppv, F);
    std::vector<const rsg_hit *> srt((size_t)nhit);
    for (int h = 0; h < nhit; h++) srt[(size_t)h] = hits + h;
    std::vector<const rsg_hit *> tmp((size_t)nhit);
    if (nhit > 1) merge_sort(srt, tmp, 0, (size_t)nhit, by_score != 0);
    fprintf(fp, "#       left_pos       right_pos        score          E-value           pvalue       substitutions      power   nseff(%%)       %% canonical_pairs\n");
    fprintf(fp, "#--------------------------------------------------------------------------------------------------------------------------------------------------\n");
    if (nhit == 0) fprintf(fp, "no significant pairs\n");
    for (int h = 0; h < nhit; h++) {
      const rsg_hit &H = *srt[(size_t)h];
      double can, ns, nsf;
      pair_canonicals(spair, alen, H.i, H.j, &can, &ns, &nsf);
      fprintf(fp, "%s\t%8d\t%8d\t%.5f\t%g\t%g\t%lld\t\t%.2f\t%.2f (%.1f%%)\t%.1f\n", hit_mark(H), pos(H.i), pos(H.j), H.sc,
              H.Eval, H.pval, (long long)H.nsubs, H.power, ns, nsf, can);
    }
  }
  fclose(fp);
  return RSG_OK;
}

extern "C" int rsg_power_SPAIR_Write(const char *path, const rsg_spair *spair, int64_t dim, int in_given, char *errbuf)
{
  FILE *fp = fopen(path, "w");
  if (!fp) return rsg_fail(errbuf, "could not open %s", path);
  if (dim < 1) { fclose(fp); return RSG_OK; }
  const int ptype = spair[0].powertype;
  double expect = 0., exp_std = 0., avgsub = 0., canonical = 0., nseff_f = 0.;
  long long nbp = 0, ncv = 0;
  fprintf(fp, in_given ? "\n# Power analysis of given structure \n#\n" : "\n# Power analysis of CaCoFold structure \n#\n");
  switch (ptype) {
  case RSG_SINGLE_SUBS:
    fprintf(fp, "# Powertype = SINGLE SUBSTITUTIONS\n");
    fprintf(fp, "# covary  left_pos      right_pos    substitutions      power   nseff(%%)       %% canonical_pairs\n");
    fprintf(fp, "#-------------------------------------------------------------------------------------------------\n");
    break;
  case RSG_JOIN_SUBS:
    fprintf(fp, "# Powertype = JOIN SUBSTITUTIONS\n");
    fprintf(fp, "# covary  left_pos      right_pos    subs_join     power_join\n");
    fprintf(fp, "#----------------------------------------------------------------\n");
    break;
  default:
    fprintf(fp, "# Powertype = DOUBLE SUBSTITUTIONS\n");
    fprintf(fp, "# covary  left_pos      right_pos    subs_double     power_double\n");
    fprintf(fp, "#----------------------------------------------------------------\n");
    break;
  }
  for (int64_t n = 0; n < dim; n++) {
    const rsg_spair &s = spair[n];
    if (s.powertype != ptype) { fclose(fp); return rsg_fail(errbuf, "all pairs should use the same powertype"); }
    if (!(in_given && s.bptype_given == RSG_BP_WWC)) continue;   // the CaCoFold structure is not kept here
    nbp++;
    const char *mk = s.covary ? "     *    " : "          ";
    if (s.covary) ncv++;
    if (ptype == RSG_SINGLE_SUBS) {
      expect += s.power; exp_std += s.power * (1.0 - s.power); avgsub += (double)s.nsubs;
      canonical = 100. * (s.wcf_f + s.wobble_f);
      nseff_f = (s.nseff + s.ngap) ? 100. * (s.nseff / (s.nseff + s.ngap)) : 0.;
      fprintf(fp, "%s%lld\t\t%lld\t\t%lld\t\t%.2f\t%.2f (%.1f%%)\t%.1f\n", mk, (long long)s.iabs, (long long)s.jabs,
              (long long)s.nsubs, s.power, s.nseff, nseff_f, canonical);
    } else if (ptype == RSG_JOIN_SUBS) {
      expect += s.power_join; exp_std += s.power_join * (1.0 - s.power_join); avgsub += (double)s.nsubs_join;
      fprintf(fp, "%s%lld\t\t%lld\t\t%lld\t\t%.2f\n", mk, (long long)s.iabs, (long long)s.jabs, (long long)s.nsubs_join,
              s.power_join);
    } else {
      expect += s.power_double; exp_std += s.power_double * (1.0 - s.power_double); avgsub += (double)s.nsubs_double;
      fprintf(fp, "%s%lld\t\t%lld\t\t%lld\t\t%.2f\n", mk, (long long)s.iabs, (long long)s.jabs, (long long)s.nsubs_double,
              s.power_double);
    }
  }
  avgsub /= (nbp > 0) ? (double)nbp : 1.;
  if (exp_std > 0) exp_std = sqrt(exp_std);
  fprintf(fp, "#\n# BPAIRS %lld\n", nbp);
  fprintf(fp, "# avg substitutions per BP  %.1f\n", avgsub);
  fprintf(fp, "# BPAIRS expected to covary %.1f +/- %.1f\n", expect, exp_std);
  fprintf(fp, "# BPAIRS observed to covary %lld\n#\n", ncv);
  fclose(fp);
  return RSG_OK;
}

namespace {
// cov_histogram_PlotSurvival (src/covariation.c:1637-1673)
void plot_survival(FILE *fp, const rsg_hist *h, const double *survfit)
{
  if (h->imax >= 0 && h->obs[h->imax] > 1) fprintf(fp, "%f\t%g\n", h->xmax, 1.0 / (double)h->Nc);
  uint64_t c = 0;
  for (int i = h->imax; i >= h->imin && i >= 0; i--)
    if (h->obs[i] > 0) {
      c += h->obs[i];
      fprintf(fp, "%f\t%g\n", h->w * i + h->bmin, (double)c / (double)h->Nc);
    }
  fprintf(fp, "&\n");
  if (survfit) {
    for (int i = 2 * h->nb - 1; i >= 0; i--)
      if (survfit[i] > 0.) fprintf(fp, "%f\t%g\n", h->w * i + h->bmin, survfit[i]);
    fprintf(fp, "&\n");
  }
}
} // namespace

extern "C" int rsg_cov_WriteSurv(const char *path, const rsg_hist *hb, const rsg_hist *ht, const rsg_hist *null,
                                 const double *survfit, char *errbuf)
{
  FILE *fp = fopen(path, "w");
  if (!fp) return rsg_fail(errbuf, "could not open covhisfile %s", path);
  if (hb) plot_survival(fp, hb, nullptr);
  if (ht) plot_survival(fp, ht, nullptr);
  if (null) plot_survival(fp, null, survfit);
  fclose(fp);
  return RSG_OK;
}
