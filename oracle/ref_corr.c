/* ref_corr.c -- oracle: pair counts, probabilities, statistics, APC/ASC.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 * Serial FP64 restatement of src/correlators.c (R-scape v2.6.11) with
 * GAPASCHAR == 0 (src/rscape_config.h:48), so K = 4 and gaps are not a
 * symbol of the pair table.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "refcpu.h"

#define K   4
#define K2  16
#define IDX(x, y) ((x) * K + (y))

static int canonical(uint8_t a) { return a < K; }

/* Threads of the count loops (OpenMP over rows i; results do not depend on it). */
static int ref_threads = 1;
void ref_set_threads(int n) { ref_threads = n < 1 ? 1 : n; }
int  ref_get_threads(void) { return ref_threads; }

/* Compensated sum; restatement of Easel esl_vec_DSum (un-vendored). */
static double dsum(const double *v, int n)
{
  double s = 0., c = 0., y, t;
  for (int i = 0; i < n; i++) { y = v[i] - c; t = s + y; c = (t - s) - y; s = t; }
  return s;
}

/* Easel esl_vec_DNorm semantics: divide by the sum, uniform if the sum is 0. */
static void dnorm(double *v, int n)
{
  double s = dsum(v, n);
  if (s != 0.) for (int i = 0; i < n; i++) v[i] /= s;
  else         for (int i = 0; i < n; i++) v[i] = 1. / (double)n;
}

/* Easel esl_vec_DValidate: each entry in [0,1], sum within tol of 1. */
static int dvalidate(const double *v, int n, double tol)
{
  double s = 0.;
  for (int i = 0; i < n; i++) {
    if (!isfinite(v[i]) || v[i] < 0. || v[i] > 1.) return REF_FAIL;
    s += v[i];
  }
  return (fabs(s - 1.) > tol) ? REF_FAIL : REF_OK;
}

/* correlators.c:1695-1780 (mutual_naive_ppij) driven by corr_NaivePP :1300-1317.
 * For i<j: c[x][y] += w_s when both residues are canonical, else ngap += w_s;
 * pp = normalize(c + 1e-10); pp[j][i] is the transpose; ngap[j][i] is not written. */
int ref_naive_pp(const uint8_t *ax, const double *wgt, int64_t N, int64_t L,
                 double *pp, double *nseff, double *ngap)
{
  /* the columns are read from a column-major copy (contiguous over s); each pair still sums
     over s = 0..N-1 in the reference's order.  Rows i are independent, so the threaded loop
     (ref_set_threads; 1 thread by default) gives bit-identical results */
  uint8_t *axT = malloc((size_t)(N * L));
  if (!axT) return REF_FAIL;
  for (int64_t s = 0; s < N; s++)
    for (int64_t i = 0; i < L; i++) axT[i * N + s] = ax[s * L + i];
#pragma omp parallel for schedule(dynamic, 1) num_threads(ref_threads)
  for (int64_t i = 0; i < L - 1; i++)
    for (int64_t j = i + 1; j < L; j++) {
      const uint8_t *ci = axT + i * N, *cj = axT + j * N;
      double *p = pp + (i * L + j) * K2;
      for (int c = 0; c < K2; c++) p[c] = 1e-10;             /* :1713 */
      double ns = 0., ng = 0.;                                 /* zeroed by corr_Reuse :1221-1251 */
      for (int64_t s = 0; s < N; s++) {
        uint8_t ri = ci[s], rj = cj[s];
        if (canonical(ri) && canonical(rj)) {
          ns += wgt[s];
          p[IDX(ri, rj)] += wgt[s];
        } else {
          ng += wgt[s];                                        /* :1752 */
        }
      }
      nseff[i * L + j] = ns;
      ngap[i * L + j]  = ng;
      dnorm(p, K2);                                            /* :1758 */
      double *q = pp + (j * L + i) * K2;
      for (int x = 0; x < K; x++)
        for (int y = 0; y < K; y++) q[IDX(y, x)] = p[IDX(x, y)];
      nseff[j * L + i] = nseff[i * L + j];
    }
  free(axT);
  return REF_OK;
}

/* Unit-weight integer 4x4 counts, the quantity the int8 Gram produces. */
int ref_pair_counts(const uint8_t *ax, int64_t N, int64_t L, int32_t *cnt)
{
  memset(cnt, 0, sizeof(int32_t) * (size_t)(L * L * K2));
  uint8_t *axT = malloc((size_t)(N * L));
  if (!axT) return REF_FAIL;
  for (int64_t s = 0; s < N; s++)
    for (int64_t i = 0; i < L; i++) axT[i * N + s] = ax[s * L + i];
#pragma omp parallel for schedule(dynamic, 1) num_threads(ref_threads)
  for (int64_t i = 0; i < L - 1; i++)
    for (int64_t j = i + 1; j < L; j++) {
      const uint8_t *ci = axT + i * N, *cj = axT + j * N;
      int32_t *c = cnt + (i * L + j) * K2;
      for (int64_t s = 0; s < N; s++) {
        uint8_t ri = ci[s], rj = cj[s];
        if (canonical(ri) && canonical(rj)) c[IDX(ri, rj)]++;
      }
    }
  free(axT);
  return REF_OK;
}

/* correlators.c:1782-1817: ps[i][x] = normalize(1e-5 + sum_s w_s [a_si == x]), x in 0..4 (gap included). */
int ref_naive_ps(const uint8_t *ax, const double *wgt, int64_t N, int64_t L, double *ps)
{
  for (int64_t i = 0; i < L; i++) {
    double *p = ps + i * (K + 1);
    for (int x = 0; x <= K; x++) p[x] = 1e-5;
    for (int64_t s = 0; s < N; s++) {
      uint8_t r = ax[s * L + i];
      if (r < K + 1) p[r] += wgt[s];
    }
    dnorm(p, K + 1);
  }
  return REF_OK;
}

/* correlators.c:1337-1375: pm[i][x] = normalize_x( sum_{j: nseff[i][j]>0} sum_y pp[i][j][x][y] ).
 * nseff[i][i] stays 0, so j == i never contributes. */
int ref_marginals(int64_t L, const double *pp, const double *nseff, double tol, double *pm, char *errbuf)
{
  for (int64_t i = 0; i < L; i++) {
    double *m = pm + i * K;
    for (int x = 0; x < K; x++) m[x] = 0.;
    for (int x = 0; x < K; x++)
      for (int64_t j = 0; j < L; j++)
        for (int y = 0; y < K; y++)
          if (nseff[i * L + j] > 0) m[x] += pp[(i * L + j) * K2 + IDX(x, y)];
    dnorm(m, K);
    if (dvalidate(m, K, tol) != REF_OK) {
      if (errbuf) snprintf(errbuf, 128, "pm[%lld] validation failed", (long long)i);
      return REF_FAIL;
    }
  }
  return REF_OK;
}

/* correlators.c:1499-1545 */
int ref_validate_probs(int64_t L, const double *pp, const double *pm, const double *ps, double tol, char *errbuf)
{
  for (int64_t i = 0; i < L - 1; i++)
    for (int64_t j = i + 1; j < L; j++)
      if (dvalidate(pp + (i * L + j) * K2, K2, tol) != REF_OK) {
        if (errbuf) snprintf(errbuf, 128, "pp validation failed");
        return REF_FAIL;
      }
  for (int64_t i = 0; i < L; i++) {
    if (dvalidate(pm + i * K, K, tol) != REF_OK) { if (errbuf) snprintf(errbuf, 128, "pm validation failed"); return REF_FAIL; }
    if (dvalidate(ps + i * (K + 1), K + 1, tol) != REF_OK) { if (errbuf) snprintf(errbuf, 128, "ps validation failed"); return REF_FAIL; }
  }
  return REF_OK;
}

/* correlators.c:1423-1496, method NONPARAM. */
int ref_probs(const uint8_t *ax, const double *wgt, int64_t N, int64_t L, double tol,
              double *pp, double *nseff, double *ngap, double *pm, double *ps, char *errbuf)
{
  memset(pp, 0, sizeof(double) * (size_t)(L * L * K2));
  memset(nseff, 0, sizeof(double) * (size_t)(L * L));
  memset(ngap, 0, sizeof(double) * (size_t)(L * L));
  if (ref_naive_pp(ax, wgt, N, L, pp, nseff, ngap) != REF_OK) return REF_FAIL;
  if (ref_naive_ps(ax, wgt, N, L, ps) != REF_OK) return REF_FAIL;
  if (ref_marginals(L, pp, nseff, tol, pm, errbuf) != REF_OK) return REF_FAIL;
  return ref_validate_probs(L, pp, pm, ps, tol, errbuf);
}

/* correlators.c:1637-1642 */
static int allowed(int x, int y, const double *allowpair)
{
  return (x < 4 && y < 4 && allowpair[x * 4 + y] > 0) ? 1 : 0;
}

/* COV reset of corr_ReuseCOV (correlators.c:1254-1268). */
static void reset_cov(int64_t L, double *COV, double *minmax)
{
  for (int64_t c = 0; c < L * L; c++) COV[c] = -INFINITY;
  minmax[0] = INFINITY; minmax[1] = -INFINITY;
}

static void put(int64_t L, double *COV, double *minmax, int64_t i, int64_t j, double v)
{
  COV[i * L + j] = COV[j * L + i] = v;
  if (v < minmax[0]) minmax[0] = v;
  if (v > minmax[1]) minmax[1] = v;
}

/* One C16 pair value; CHI :96-121, OMES :230-254, GT :360-395, MI :534-567,
 * MIr :661-697, MIg :794-827.  CWC (GT only) :397-434 restricts to allowed cells. */
static double stat_c16(int type, int cwc, const double *allowpair, const double *p,
                       const double *pmi, const double *pmj, double ns, double ng)
{
  double v = 0., H = 0.;
  for (int x = 0; x < K; x++)
    for (int y = 0; y < K; y++) {
      double pxy = p[IDX(x, y)];
      double ex  = ns * pmi[x] * pmj[y];
      double ob  = ns * pxy;
      switch (type) {
      case REF_CHI:  v += (ex > 0.) ? (ob - ex) * (ob - ex) / ex : 0.0; break;
      case REF_OMES: v += (ex > 0.) ? (ob - ex) * (ob - ex) / ns : 0.0; break;
      case REF_GT:
        if (cwc && !allowed(x, y, allowpair)) break;
        v += (ex > 0. && ob > 0.) ? ob * log(ob / ex) : 0.0; break;
      case REF_MI: case REF_MIg:
        v += (pxy > 0.0 && pmi[x] > 0.0 && pmj[y] > 0.0) ? pxy * (log(pxy) - log(pmi[x]) - log(pmj[y])) : 0.0; break;
      case REF_MIr:
        H -= (pxy > 0.0) ? pxy * log(pxy) : 0.0;
        v += (pxy > 0.0 && pmi[x] > 0.0 && pmj[y] > 0.0) ? pxy * (log(pxy) - log(pmi[x]) - log(pmj[y])) : 0.0; break;
      }
    }
  if (type == REF_GT)  v *= 2.0;
  if (type == REF_MIr) v = (H > 1e-2) ? v / H : 0.0;
  if (type == REF_MIg) v -= (ns > 0) ? ng / ns : 0.0;
  return v;
}

/* C2 pooled value: CHI :146-173, OMES :278-305, GT :463-485, MI :590-606,
 * MIr :717-737, MIg :846-867.  Note MI_C2/MIg_C2 do not test q > 0. */
static double stat_c2(int type, const double *allowpair, const double *p,
                      const double *pmi, const double *pmj, double ns, double ng)
{
  double pw = 0., pn = 0., qw = 0., qn = 0., v = 0., H = 0.;
  for (int x = 0; x < K; x++)
    for (int y = 0; y < K; y++) {
      if (allowed(x, y, allowpair)) { pw += p[IDX(x, y)]; qw += pmi[x] * pmj[y]; }
      else                          { pn += p[IDX(x, y)]; qn += pmi[x] * pmj[y]; }
    }
  double ew = ns * qw, en = ns * qn, ow = ns * pw, on = ns * pn;
  switch (type) {
  case REF_CHI:
    v += (ew > 0.) ? (ow - ew) * (ow - ew) / ew : 0.0;
    v += (en > 0.) ? (on - en) * (on - en) / en : 0.0; break;
  case REF_OMES:
    v += (ew > 0.) ? (ow - ew) * (ow - ew) / ns : 0.0;
    v += (en > 0.) ? (on - en) * (on - en) / ns : 0.0; break;
  case REF_GT:
    v += (ew > 0. && ow > 0.) ? ow * log(ow / ew) : 0.0;
    v += (en > 0. && on > 0.) ? on * log(on / en) : 0.0;
    v *= 2.0; break;
  case REF_MI: case REF_MIg:
    v += (pw > 0.) ? pw * (log(pw) - log(qw)) : 0.0;
    v += (pn > 0.) ? pn * (log(pn) - log(qn)) : 0.0;
    if (type == REF_MIg) v -= (ns > 0) ? ng / ns : 0.0;
    break;
  case REF_MIr:
    H -= (pw > 0.) ? pw * log(pw) : 0.0;
    H -= (pn > 0.) ? pn * log(pn) : 0.0;
    v += (pw > 0. && qw > 0.0) ? pw * (log(pw) - log(qw)) : 0.0;
    v += (pn > 0. && qn > 0.0) ? pn * (log(pn) - log(qn)) : 0.0;
    v = (H > 1e-2) ? v / H : 0.0; break;
  }
  return v;
}

/* correlators.c:1010-1061 (CCF; C16 only, global meanp over i<j). */
static int stat_ccf(int64_t L, const double *pm, const double *nseff, double *COV, double *minmax)
{
  double meanp[K];
  for (int x = 0; x < K; x++) {
    meanp[x] = 0.0;
    for (int64_t i = 0; i < L; i++)
      for (int64_t j = i + 1; j < L; j++) meanp[x] += nseff[i * L + j] * pm[i * K + x];
  }
  dnorm(meanp, K);
  for (int64_t i = 0; i < L - 1; i++)
    for (int64_t j = i + 1; j < L; j++) {
      double ccf = 0.0, ns = nseff[i * L + j];
      for (int x = 0; x < K; x++)
        for (int y = 0; y < K; y++) {
          double cc = (ns * pm[i * K + x] - meanp[x]) * (ns * pm[j * K + y] - meanp[y]);
          ccf += cc * cc;
        }
      put(L, COV, minmax, i, j, sqrt(ccf));
    }
  return REF_OK;
}

/* Dispatch of corr_Calculate{CHI,OMES,GT,MI,MIr,MIg,CCF} with the CSELECT rule
 * (correlators.c:59-70, :335-341): C2 if nseq <= nseqthresh or alen <= alenthresh. */
int ref_cov_stat(int covtype, int covclass, int64_t nseq, int64_t L, int nseqthresh, int alenthresh,
                 const double *allowpair, const double *pp, const double *pm,
                 const double *nseff, const double *ngap, double *COV, double *minmax, char *errbuf)
{
  reset_cov(L, COV, minmax);
  if (covtype == REF_CCF) return stat_ccf(L, pm, nseff, COV, minmax);
  int cls = covclass;
  if (cls == REF_CSELECT) cls = (nseq <= nseqthresh || L <= alenthresh) ? REF_C2 : REF_C16;
  if (cls == REF_CWC && covtype != REF_GT) {
    if (errbuf) snprintf(errbuf, 128, "CWC not implemented for this statistic");
    return REF_FAIL;
  }
  for (int64_t i = 0; i < L - 1; i++)
    for (int64_t j = i + 1; j < L; j++) {
      const double *p = pp + (i * L + j) * K2;
      double ns = nseff[i * L + j], ng = ngap[i * L + j], v;
      if (cls == REF_C2) v = stat_c2(covtype, allowpair, p, pm + i * K, pm + j * K, ns, ng);
      else               v = stat_c16(covtype, cls == REF_CWC, allowpair, p, pm + i * K, pm + j * K, ns, ng);
      put(L, COV, minmax, i, j, v);
    }
  return REF_OK;
}

/* correlators.c:876-933 (RAF; unweighted, literal O(L^2 N^2) loop). */
int ref_raf(const uint8_t *ax, int64_t N, int64_t L, const double *allowpair, double *COV, double *minmax)
{
  reset_cov(L, COV, minmax);
  for (int64_t i = 0; i < L - 1; i++)
    for (int64_t j = i + 1; j < L; j++) {
      double cij = 0., qij = 0.;
      for (int64_t s1 = 0; s1 < N; s1++) {
        int ai = ax[s1 * L + i], aj = ax[s1 * L + j];
        int a_ok = allowed(ai, aj, allowpair);
        if (!a_ok) qij += 1.0;
        for (int64_t s2 = s1 + 1; s2 < N; s2++) {
          int bi = ax[s2 * L + i], bj = ax[s2 * L + j];
          if (a_ok && allowed(bi, bj, allowpair)) {
            if (ai != bi && aj != bj)      cij += 2.0;
            else if (ai != bi || aj != bj) cij += 1.0;
          }
        }
      }
      qij /= (double)N;
      cij /= (N > 1) ? (double)N * ((double)N - 1.0) : 1.0;
      cij *= 2.0;
      put(L, COV, minmax, i, j, cij - qij);
    }
  return REF_OK;
}

/* RAF of correlators.c:876-933 from the unweighted 4x4 pair counts instead of the O(N^2)
 * sequence-pair loop (for alignments where that loop cannot finish).  With n_xy the count of
 * allowed pairs (x, y) at (i, j), M = sum n_xy, m_x / m'_y its row / column sums:
 *   sum_{s1<s2} A_s1 A_s2 [a_s1 != a_s2 at i] = (M^2 - sum_x m_x^2) / 2, same at j, and a pair of
 *   sequences scores [differ at i] + [differ at j] (2 when both, 1 when one), so
 *   cij = 2 / (N (N - 1)) ((M^2 - sum m_x^2) / 2 + (M^2 - sum m'_y^2) / 2), qij = (N - M) / N.
 * Pinned against ref_raf (the literal loop) on small alignments. */
int ref_raf_counts(const int32_t *cnt, int64_t N, int64_t L, const double *allowpair, double *COV, double *minmax)
{
  reset_cov(L, COV, minmax);
  for (int64_t i = 0; i < L - 1; i++)
    for (int64_t j = i + 1; j < L; j++) {
      const int32_t *c = cnt + (i * L + j) * K2;
      double M = 0., mx[K] = {0., 0., 0., 0.}, my[K] = {0., 0., 0., 0.};
      for (int x = 0; x < K; x++)
        for (int y = 0; y < K; y++)
          if (allowed(x, y, allowpair)) { double n = c[IDX(x, y)]; M += n; mx[x] += n; my[y] += n; }
      double sx = 0., sy = 0.;
      for (int k = 0; k < K; k++) { sx += mx[k] * mx[k]; sy += my[k] * my[k]; }
      double cij = 0.5 * (M * M - sx) + 0.5 * (M * M - sy);
      cij /= (N > 1) ? (double)N * ((double)N - 1.0) : 1.0;
      cij *= 2.0;
      put(L, COV, minmax, i, j, cij - ((double)N - M) / (double)N);
    }
  return REF_OK;
}

/* The RAFS stencil (correlators.c:954-960) applied to a RAF matrix B (both triangles). */
int ref_rafs_from(const double *B, int64_t L, double *COV, double *minmax)
{
  reset_cov(L, COV, minmax);
  for (int64_t i = 0; i < L - 1; i++)
    for (int64_t j = i + 1; j < L; j++) {
      double v = 2.0 * B[i * L + j];
      if (i > 0 && j < L - 1)               v += B[(i - 1) * L + (j + 1)];
      if (j > 0 && i < L - 1 && i < j - 2)  v += B[(i + 1) * L + (j - 1)];
      put(L, COV, minmax, i, j, 0.25 * v);
    }
  return REF_OK;
}

/* correlators.c:935-982 (RAFS: 3-point anti-diagonal stencil over RAF, always x0.25). */
int ref_rafs(const uint8_t *ax, int64_t N, int64_t L, const double *allowpair, double *COV, double *minmax)
{
  double *b = malloc(sizeof(double) * (size_t)(L * L));
  if (!b) return REF_FAIL;
  ref_raf(ax, N, L, allowpair, COV, minmax);
  memcpy(b, COV, sizeof(double) * (size_t)(L * L));
  reset_cov(L, COV, minmax);
  for (int64_t i = 0; i < L - 1; i++)
    for (int64_t j = i + 1; j < L; j++) {
      double v = 2.0 * b[i * L + j];
      if (i > 0 && j < L - 1)               v += b[(i - 1) * L + (j + 1)];
      if (j > 0 && i < L - 1 && i < j - 2)  v += b[(i + 1) * L + (j - 1)];
      put(L, COV, minmax, i, j, 0.25 * v);
    }
  free(b);
  return REF_OK;
}

/* correlators.c:1063-1157.  COVavg over i<j, COVx over j != i, then
 * APC: COV - COVx_i COVx_j / COVavg (0 if COVavg == 0); ASC: COV - (COVx_i + COVx_j - COVavg). */
int ref_cov_corrected(int actype, int64_t L, double *COV, double *minmax, int shiftnonneg, char *errbuf)
{
  double *C = malloc(sizeof(double) * (size_t)(L * L));
  double *X = malloc(sizeof(double) * (size_t)L);
  double avg = 0.;
  if (!C || !X) { free(C); free(X); return REF_FAIL; }
  memcpy(C, COV, sizeof(double) * (size_t)(L * L));
  reset_cov(L, COV, minmax);
  for (int64_t i = 0; i < L - 1; i++)
    for (int64_t j = i + 1; j < L; j++) avg += C[i * L + j];
  if (L > 1) avg /= (double)L * ((double)L - 1.);
  avg *= 2.;
  for (int64_t i = 0; i < L; i++) {
    X[i] = 0.0;
    for (int64_t j = 0; j < L; j++) if (j != i) X[i] += C[i * L + j];
    if (L > 1) X[i] /= (double)L - 1.;
  }
  for (int64_t i = 0; i < L; i++)
    for (int64_t j = 0; j < L; j++) {
      if (i == j) continue;
      double v;
      if (actype == REF_APC) v = (avg != 0.0) ? C[i * L + j] - X[i] * X[j] / avg : 0.0;
      else                   v = C[i * L + j] - (X[i] + X[j] - avg);
      if (isnan(v)) {
        if (errbuf) snprintf(errbuf, 128, "bad covariation COVavg %f", avg);
        free(C); free(X); return REF_FAIL;
      }
      COV[i * L + j] = v;
      if (v < minmax[0]) minmax[0] = v;
      if (v > minmax[1]) minmax[1] = v;
    }
  if (shiftnonneg)
    for (int64_t i = 0; i < L; i++)
      for (int64_t j = 0; j < L; j++) if (i != j) COV[i * L + j] -= minmax[0];
  free(C); free(X);
  return REF_OK;
}
