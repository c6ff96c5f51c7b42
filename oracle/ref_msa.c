/* ref_msa.c -- oracle: MSA preprocessing and sequence weights.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Gap-column filter and symbol canonicalisation restate src/msamanip.c.
 * Sequence weights live in Easel (un-vendored, version not pinned):
 *   - GSC (Gerstein, Sonnhammer & Chothia 1994) on a UPGMA tree built from the
 *     fractional-difference matrix D = 1 - idents/min(len1,len2) over canonical
 *     residues, weights renormalised to sum to N.  Pinned end-to-end by the
 *     tutorial GTp scores (documentation/tutorial.tex:193-212).
 *   - PB (Henikoff & Henikoff 1994) position-based weights: PARITY UNPINNED
 *     (pinned only by the published formula and hand examples).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "refcpu.h"

static int is_residue(uint8_t a)   { return a < 4 || (a >= 5 && a <= 15); }  /* canonical, degenerate or N */
static int is_gap(uint8_t a)       { return a == 4; }
static int is_missing(uint8_t a)   { return a == 17; }

/* msamanip.c:461-541: keep column if r > 0 and r / totwgt >= 1 - gapthresh, where r
 * sums weights of residues and totwgt sums residues + gaps (missing ignored). */
int ref_gapcol_filter(const uint8_t *ax, const double *wgt, int64_t N, int64_t L, double gapthresh, int *useme)
{
  double idthresh = 1.0 - gapthresh;
  int kept = 0;
  for (int64_t c = 0; c < L; c++) {
    double r = 0., tot = 0.;
    for (int64_t i = 0; i < N; i++) {
      uint8_t a = ax[i * L + c];
      if (is_residue(a))      { r += wgt[i]; tot += wgt[i]; }
      else if (is_gap(a))     tot += wgt[i];
      else if (is_missing(a)) continue;
    }
    useme[c] = (r > 0. && r / tot >= idthresh) ? 1 : 0;
    kept += useme[c];
  }
  return kept;
}

/* msamanip.c:284-296 (degenerate -> N) and :315-330 (missing / nonresidue -> gap). */
void ref_canonicalize(uint8_t *ax, int64_t n)
{
  for (int64_t k = 0; k < n; k++) {
    uint8_t a = ax[k];
    if (a >= 5 && a <= 14) ax[k] = 15;
    else if (a == 16 || a == 17) ax[k] = 4;
  }
}

int ref_weights_pb(const uint8_t *ax, int64_t N, int64_t L, double *wgt)
{
  int nres[4];
  for (int64_t s = 0; s < N; s++) wgt[s] = 0.;
  for (int64_t c = 0; c < L; c++) {
    memset(nres, 0, sizeof(nres));
    for (int64_t s = 0; s < N; s++) if (ax[s * L + c] < 4) nres[ax[s * L + c]]++;
    int ntot = 0;
    for (int x = 0; x < 4; x++) if (nres[x] > 0) ntot++;
    if (ntot == 0) continue;
    for (int64_t s = 0; s < N; s++)
      if (ax[s * L + c] < 4) wgt[s] += 1. / (double)(ntot * nres[ax[s * L + c]]);
  }
  for (int64_t s = 0; s < N; s++) {
    int64_t rlen = 0;
    for (int64_t c = 0; c < L; c++) if (ax[s * L + c] < 4) rlen++;
    if (rlen > 0) wgt[s] /= (double)rlen;
  }
  double sum = 0.;
  for (int64_t s = 0; s < N; s++) sum += wgt[s];
  if (sum > 0.) for (int64_t s = 0; s < N; s++) wgt[s] *= (double)N / sum;
  else          for (int64_t s = 0; s < N; s++) wgt[s] = 1.;
  return REF_OK;
}

/* UPGMA clustering into a rooted binary tree: internal nodes 0..N-2 (root 0),
 * children > 0 are internal nodes, children <= 0 are leaves -child. */
static int upgma(double *D, int N, int *left, int *right, double *ld, double *rd)
{
  int    *idx = malloc(sizeof(int) * N), *nin = malloc(sizeof(int) * N);
  double *height = calloc((size_t)(N > 1 ? N - 1 : 1), sizeof(double));
  int n = N, i = 0, j = 0;
  for (int k = 0; k < N; k++) { idx[k] = -k; nin[k] = 1; }
  while (n >= 2) {
    double minD = INFINITY;
    for (int r = 0; r < n; r++)
      for (int c = r + 1; c < n; c++)
        if (D[r * N + c] < minD) { minD = D[r * N + c]; i = r; j = c; }
    int v = n - 2;
    left[v] = idx[i]; right[v] = idx[j];
    height[v] = minD / 2.;
    ld[v] = rd[v] = height[v];
    if (idx[i] > 0) { ld[v] -= height[idx[i]]; if (ld[v] < 0.) ld[v] = 0.; }
    if (idx[j] > 0) { rd[v] -= height[idx[j]]; if (rd[v] < 0.) rd[v] = 0.; }
    /* the merged cluster takes row j; row n-1 then fills the freed row i.  Of the
     * bookkeeping orders tried, this one reproduces the tutorial's published GTp
     * scores (documentation/tutorial.tex:202-212) most closely. */
    for (int c = 0; c < n; c++) {
      if (c == i || c == j) continue;
      D[j * N + c] = (nin[j] * D[j * N + c] + nin[i] * D[i * N + c]) / (double)(nin[i] + nin[j]);
      D[c * N + j] = D[j * N + c];
    }
    idx[j] = v; nin[j] += nin[i];
    for (int c = 0; c < n; c++) { D[i * N + c] = D[(n - 1) * N + c]; D[c * N + i] = D[c * N + (n - 1)]; }
    idx[i] = idx[n - 1]; nin[i] = nin[n - 1];
    n--;
  }
  free(idx); free(nin); free(height);
  return REF_OK;
}

int ref_weights_gsc(const uint8_t *ax, int64_t N, int64_t L, double *wgt)
{
  if (N == 1) { wgt[0] = 1.0; return REF_OK; }
  double *D = malloc(sizeof(double) * (size_t)(N * N));
  for (int64_t a = 0; a < N; a++) {
    D[a * N + a] = 0.;
    for (int64_t b = a + 1; b < N; b++) {
      int64_t id = 0, l1 = 0, l2 = 0;
      for (int64_t c = 0; c < L; c++) {
        uint8_t x = ax[a * L + c], y = ax[b * L + c];
        if (x < 4) l1++;
        if (y < 4) l2++;
        if (x < 4 && y < 4 && x == y) id++;
      }
      int64_t len = (l2 < l1) ? l2 : l1;
      double pid = (len == 0) ? 0. : (double)id / (double)len;
      D[a * N + b] = D[b * N + a] = 1. - pid;
    }
  }
  int nn = (int)N;
  int *left = malloc(sizeof(int) * nn), *right = malloc(sizeof(int) * nn);
  double *ld = malloc(sizeof(double) * nn), *rd = malloc(sizeof(double) * nn), *x = malloc(sizeof(double) * nn);
  upgma(D, nn, left, right, ld, rd);
  /* total branch length under each internal node (children have larger indices) */
  for (int v = nn - 2; v >= 0; v--) {
    x[v] = ld[v] + rd[v];
    if (left[v] > 0)  x[v] += x[left[v]];
    if (right[v] > 0) x[v] += x[right[v]];
  }
  /* top-down: distribute the weight carried by the edge above each node */
  for (int v = 0; v <= nn - 2; v++) {
    if (v == 0) x[v] = 0.;
    double lw = ld[v] + ((left[v] > 0) ? x[left[v]] : 0.);
    double rw = rd[v] + ((right[v] > 0) ? x[right[v]] : 0.);
    double lx, rx;
    if (lw + rw == 0.) { lx = rx = 0.5 * x[v]; }
    else { lx = x[v] * lw / (lw + rw); rx = x[v] * rw / (lw + rw); }
    if (left[v] <= 0) wgt[-left[v]] = lx + ld[v]; else x[left[v]] = lx + ld[v];
    if (right[v] <= 0) wgt[-right[v]] = rx + rd[v]; else x[right[v]] = rx + rd[v];
  }
  double s = 0., cmp = 0.;
  for (int64_t k = 0; k < N; k++) { double yy = wgt[k] - cmp, t = s + yy; cmp = (t - s) - yy; s = t; }
  if (s != 0.) for (int64_t k = 0; k < N; k++) wgt[k] = wgt[k] / s * (double)N;
  else         for (int64_t k = 0; k < N; k++) wgt[k] = 1.;
  free(D); free(left); free(right); free(ld); free(rd); free(x);
  return REF_OK;
}

/* R-scape.c:2447-2456 (the code's defaults, which differ from the user guide). */
int ref_nshuffle_default(int64_t nseq, int64_t alen)
{
  int ns = 20;
  if (nseq * alen < 1e3) ns = 200;
  if (alen < 50)         ns = 200;
  if (alen < 40)         ns = 300;
  if (nseq < 4)          ns = 300;
  return ns;
}
