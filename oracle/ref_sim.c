/* ref_sim.c -- oracle: counter-based RNG, rate-matrix evolution down a tree,
 * and R-scape's phylogenetic null (Fitch ancestors + column shuffle + replay).
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * The reference draws from Easel's sequential RNG (un-vendored); that stream
 * cannot be reproduced.  This restatement replaces it with Philox4x32-10
 * (Salmon, Moraes, Dror & Shaw, SC'11), keyed so that every draw is a pure
 * function of (seed, stream, node, position, purpose).  The GPU simulator uses
 * the same keys, so the two agree symbol for symbol; agreement with the
 * reference itself is distributional only.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "refcpu.h"

static void mulhilo(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo)
{
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32); *lo = (uint32_t)p;
}

void ref_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; r++) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(0xD2511F53u, c0, &hi0, &lo0);
    mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* One 32-bit draw for (seed, stream, node, pos); purpose tag 0. */
uint32_t ref_rng_u32(uint64_t seed, uint32_t stream, uint32_t node, uint32_t pos)
{
  uint32_t ctr[4] = { pos, node, stream, 0u }, key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) }, o[4];
  ref_philox4x32(ctr, key, o);
  return o[0];
}

static uint32_t rng_tag(uint64_t seed, uint32_t stream, uint32_t node, uint32_t pos, uint32_t tag)
{
  uint32_t ctr[4] = { pos, node, stream, tag }, key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) }, o[4];
  ref_philox4x32(ctr, key, o);
  return o[0];
}

/* The per-node block of the GTR evolution (tag 5, keyed by the internal node v and the column):
 * words 0 / 1 are the left / right child's symbol draws, 2 / 3 their leaf-gap draws. */
static uint32_t node_word(uint64_t seed, uint32_t stream, uint32_t v, uint32_t pos, int k)
{
  uint32_t ctr[4] = { pos, v, stream, 5u }, key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) }, o[4];
  ref_philox4x32(ctr, key, o);
  return o[k];
}

/* Inverse CDF over n categories with n-1 uint32 thresholds: the reference's
 * "pdf += P[i]; if (pdf > x) break" (cov_simulate.c:756-763) with x = u / 2^32. */
static int icdf(uint32_t u, const uint32_t *cdf, int n)
{
  int k = 0;
  for (int t = 0; t < n - 1; t++) k += (u >= cdf[t]);
  return k;
}

/* P(t) row a as cumulative uint32 thresholds from the eigendecomposed model
 * (model[0..15] = Pi^-1/2 V, [16..31] = V^T Pi^1/2, [32..35] = eigenvalues, [36..39] = pi). */
static void row_cdf(const double *model, double t, int a, uint32_t *out)
{
  double e[4], cum = 0.;
  for (int k = 0; k < 4; k++) e[k] = exp(model[32 + k] * t);
  for (int b = 0; b < 3; b++) {
    double p = 0.;
    for (int k = 0; k < 4; k++) p += model[a * 4 + k] * e[k] * model[16 + k * 4 + b];
    cum += (p > 0.) ? p : 0.;
    double v = cum * 4294967296.0;
    out[b] = (v >= 4294967295.0) ? 0xFFFFFFFFu : (uint32_t)v;
  }
}

/* cov_simulate.c:288-324 + :387-451 + :585-631 + :724-798 (ungapped evolution down a tree).
 * Nodes: internal v = 0..nleaf-2 (root 0, parents before children) are rows 0..nleaf-2;
 * leaf l is row nleaf-1+l.  left/right > 0 are internal, <= 0 leaves.  The root is drawn
 * from pi with the key of row 2N-1; each child symbol from P(t) of its branch; a leaf
 * symbol becomes a gap when the tag-1 draw is below gapthr.  leaves: N x L output. */
int ref_evolve_gtr(int64_t nleaf, int64_t L, const int *left, const int *right, const double *ld,
                   const double *rd, const double *model, uint64_t seed, uint32_t stream, uint32_t gapthr,
                   const uint8_t *root, uint8_t *leaves)
{
  int64_t nrow = 2 * nleaf - 1;
  uint8_t *S = malloc((size_t)(nrow * L));
  uint32_t pic[3], tab[12];
  double cum = 0.;
  for (int b = 0; b < 3; b++) {
    cum += model[36 + b];
    double v = cum * 4294967296.0;
    pic[b] = (v >= 4294967295.0) ? 0xFFFFFFFFu : (uint32_t)v;
  }
  for (int64_t pos = 0; pos < L; pos++) {
    uint32_t u = ref_rng_u32(seed, stream, (uint32_t)nrow, (uint32_t)pos);
    int a = 3;
    for (int b = 0; b < 3; b++) if (u < pic[b]) { a = b; break; }
    S[pos] = (uint8_t)a;
  }
  if (root) memcpy(S, root, (size_t)L);            /* a given root row (cov_add_root, :204-238) */
  for (int64_t v = 0; v < nleaf - 1; v++)
    for (int side = 0; side < 2; side++) {
      int d = side ? right[v] : left[v];
      int64_t row = (d > 0) ? d : (nleaf - 1 - d);
      double t = side ? rd[v] : ld[v];
      for (int a = 0; a < 4; a++) row_cdf(model, t, a, tab + a * 3);
      for (int64_t pos = 0; pos < L; pos++) {
        uint32_t u = node_word(seed, stream, (uint32_t)v, (uint32_t)pos, side);
        int sym = icdf(u, tab + S[v * L + pos] * 3, 4);
        S[row * L + pos] = (uint8_t)sym;
        if (d <= 0) {
          if (gapthr && node_word(seed, stream, (uint32_t)v, (uint32_t)pos, 2 + side) < gapthr) sym = 4;
          leaves[(-d) * L + pos] = (uint8_t)sym;
        }
      }
    }
  free(S);
  return REF_OK;
}

/* Structured null (cov_simulate.c:393-412, 774-798 with a structure, noss off): a base-paired
 * column pair (i, ct[i]) evolves as one 16-state character, pair state a * 4 + b; the root pair
 * is drawn from the stationary pair frequencies, a child pair from its parent pair's P(t) row.
 * m16: [0..255] Pi^-1/2 V, [256..511] V^T Pi^1/2, [512..527] eigenvalues, [528..543] pair
 * frequencies.  Keys: (row, 5' column, tag 2); leaf gaps reuse the per-node gap words.  Unpaired
 * columns (and, before they are overwritten, paired ones) evolve as in ref_evolve_gtr.
 * ct[c] = partner column or -1. */
static void row_cdf16(const double *m, double t, int a, uint32_t *out)
{
  double e[16], cum = 0.;
  for (int k = 0; k < 16; k++) e[k] = exp(m[512 + k] * t);
  for (int b = 0; b < 15; b++) {
    double p = 0.;
    for (int k = 0; k < 16; k++) p += m[a * 16 + k] * e[k] * m[256 + k * 16 + b];
    cum += (p > 0.) ? p : 0.;
    double v = cum * 4294967296.0;
    out[b] = (v >= 4294967295.0) ? 0xFFFFFFFFu : (uint32_t)v;
  }
}

int ref_evolve_pairs(int64_t nleaf, int64_t L, const int *left, const int *right, const double *ld,
                     const double *rd, const double *model, const double *m16, const int *ct, uint64_t seed,
                     uint32_t stream, uint32_t gapthr, const uint8_t *root, uint8_t *leaves)
{
  int64_t nrow = 2 * nleaf - 1;
  uint8_t *S = malloc((size_t)(nrow * L));
  uint32_t pic[3], pth[15], tab[12], tab16[16 * 15];
  double cum = 0.;
  for (int b = 0; b < 3; b++) {
    cum += model[36 + b];
    double v = cum * 4294967296.0;
    pic[b] = (v >= 4294967295.0) ? 0xFFFFFFFFu : (uint32_t)v;
  }
  cum = 0.;
  for (int b = 0; b < 15; b++) {
    cum += m16[528 + b];
    double v = cum * 4294967296.0;
    pth[b] = (v >= 4294967295.0) ? 0xFFFFFFFFu : (uint32_t)v;
  }
  for (int64_t pos = 0; pos < L; pos++) {
    uint32_t u = ref_rng_u32(seed, stream, (uint32_t)nrow, (uint32_t)pos);
    int a = 3;
    for (int b = 0; b < 3; b++) if (u < pic[b]) { a = b; break; }
    S[pos] = (uint8_t)a;
  }
  for (int64_t pos = 0; pos < L; pos++) {
    if (ct[pos] <= pos) continue;
    int st = icdf(rng_tag(seed, stream, (uint32_t)nrow, (uint32_t)pos, 2u), pth, 16);
    S[pos] = (uint8_t)(st >> 2); S[ct[pos]] = (uint8_t)(st & 3);
  }
  if (root) memcpy(S, root, (size_t)L);
  for (int64_t v = 0; v < nleaf - 1; v++)
    for (int side = 0; side < 2; side++) {
      int d = side ? right[v] : left[v];
      int64_t row = (d > 0) ? d : (nleaf - 1 - d);
      double t = side ? rd[v] : ld[v];
      for (int a = 0; a < 4; a++) row_cdf(model, t, a, tab + a * 3);
      for (int a = 0; a < 16; a++) row_cdf16(m16, t, a, tab16 + a * 15);
      for (int64_t pos = 0; pos < L; pos++) {
        uint32_t u = node_word(seed, stream, (uint32_t)v, (uint32_t)pos, side);
        int sym = icdf(u, tab + S[v * L + pos] * 3, 4);
        S[row * L + pos] = (uint8_t)sym;
        if (d <= 0) {
          if (gapthr && node_word(seed, stream, (uint32_t)v, (uint32_t)pos, 2 + side) < gapthr) sym = 4;
          leaves[(-d) * L + pos] = (uint8_t)sym;
        }
      }
      for (int64_t pos = 0; pos < L; pos++) {
        int64_t q = ct[pos];
        if (q <= pos) continue;
        int a = 4 * S[v * L + pos] + S[v * L + q];
        int st = icdf(rng_tag(seed, stream, (uint32_t)row, (uint32_t)pos, 2u), tab16 + a * 15, 16);
        int x = st >> 2, y = st & 3;
        S[row * L + pos] = (uint8_t)x; S[row * L + q] = (uint8_t)y;
        if (d <= 0) {
          if (gapthr && node_word(seed, stream, (uint32_t)v, (uint32_t)pos, 2 + side) < gapthr) x = 4;
          if (gapthr && node_word(seed, stream, (uint32_t)v, (uint32_t)q, 2 + side) < gapthr) y = 4;
          leaves[(-d) * L + pos] = (uint8_t)x; leaves[(-d) * L + q] = (uint8_t)y;
        }
      }
    }
  free(S);
  return REF_OK;
}

/* R-scape's null (R-scape.c:1650-1697): Fitch ancestral reconstruction per column
 * (msatree.c:173-229, 1700-1931), one shuffle of all columns of leaves+ancestors
 * (msamanip.c:1461 -> 1164), then for every branch parent -> child replay the
 * original branch's (old -> new) substitution counts at random positions of the
 * shuffled parent holding old (msamanip.c:1597-1780).  Symbols 0..4 (gap = 4).
 * Tree layout as ref_evolve_tree; allmsa rows follow the reference: leaves 0..N-1,
 * internal node v at N+v.  out receives the N shuffled leaves (N x L). */
#define DIM 6   /* residues 0..3, gap 4, "set" flag 5 */
int ref_fitch_null_rows(const uint8_t *ax, int64_t N, int64_t L, const int *left, const int *right,
                        uint64_t seed, uint8_t *all_out, uint8_t *sh_out);

int ref_fitch_null(const uint8_t *ax, int64_t N, int64_t L, const int *left, const int *right,
                   uint64_t seed, uint8_t *out)
{
  uint8_t *sh = malloc((size_t)((2 * N - 1) * L));
  int st = ref_fitch_null_rows(ax, N, L, left, right, seed, NULL, sh);
  memcpy(out, sh, (size_t)(N * L));
  free(sh);
  return st;
}

/* As ref_fitch_null, returning every row: all_out = the Fitch reconstruction (leaves 0..N-1,
 * internal node v at N+v) before the shuffle, sh_out = the shuffled, replayed alignment. */
int ref_fitch_null_rows(const uint8_t *ax, int64_t N, int64_t L, const int *left, const int *right,
                        uint64_t seed, uint8_t *all_out, uint8_t *sh_out)
{
  int64_t nn = 2 * N - 1;
  uint8_t *all = malloc((size_t)(nn * L)), *sh = malloc((size_t)(nn * L));
  int (*S)[DIM] = malloc(sizeof(int[DIM]) * (size_t)nn);
  int *parent = malloc(sizeof(int) * (size_t)N), *perm = malloc(sizeof(int) * (size_t)L);
  int *cols = malloc(sizeof(int) * (size_t)L), *cperm = malloc(sizeof(int) * (size_t)L);
  memcpy(all, ax, (size_t)(N * L));
  parent[0] = 0;
  for (int v = 0; v < N - 1; v++) {
    if (left[v] > 0)  parent[left[v]] = v;
    if (right[v] > 0) parent[right[v]] = v;
  }
  uint32_t draw = 0;
  for (int64_t c = 0; c < L; c++) {
    memset(S, 0, sizeof(int[DIM]) * (size_t)nn);
    for (int64_t n = 0; n < N; n++) {
      uint8_t a = all[n * L + c];
      if (a <= 4) S[n][a] = 1;
      else {                                              /* unknown: random symbol 0..4 */
        uint32_t u = rng_tag(seed, 1u, (uint32_t)n, (uint32_t)c, 1u);
        S[n][(int)(((uint64_t)u * 5u) >> 32)] = 1;
      }
      S[n][DIM - 1] = 1;
    }
    /* upward pass in post-order: children have larger indices than parents */
    for (int v = (int)N - 2; v >= 0; v--) {
      int il = (left[v] <= 0) ? -left[v] : (int)N + left[v];
      int ir = (right[v] <= 0) ? -right[v] : (int)N + right[v];
      int *Sv = S[N + v], *Sl = S[il], *Sr = S[ir], any = 0;
      for (int i = 0; i < DIM - 1; i++) if (Sl[i] && Sr[i]) { Sv[i] = 1; any = 1; }
      if (!any) {
        for (int i = 0; i < DIM - 2; i++) if (Sl[i] || Sr[i]) Sv[i] = 1;
        if (Sl[DIM - 2]) for (int i = 0; i < DIM - 2; i++) if (Sr[i]) Sv[i] = 1;
        if (Sr[DIM - 2]) for (int i = 0; i < DIM - 2; i++) if (Sl[i]) Sv[i] = 1;
      }
      Sv[DIM - 1] = 1;
    }
    /* choose: uniform among the set members (tree_fitch_choose with frq == NULL) */
    for (int v = 0; v < N - 1; v++) {
      int *Sv = S[N + v];
      if (v > 0) {
        int pa = all[(N + parent[v]) * L + c];
        if (Sv[pa]) for (int i = 0; i < DIM - 1; i++) if (i != pa) Sv[i] = 0;
      }
      int m = 0, opts[5];
      for (int i = 0; i < DIM - 1; i++) if (Sv[i]) opts[m++] = i;
      uint32_t u = rng_tag(seed, 2u, (uint32_t)v, (uint32_t)c, 2u);
      all[(N + v) * L + c] = (uint8_t)opts[((uint64_t)u * (uint64_t)m) >> 32];
    }
  }
  /* one permutation of all columns (Fisher-Yates, as esl_vec_IShuffle) */
  for (int64_t c = 0; c < L; c++) perm[c] = (int)c;
  for (int64_t n = L; n > 1; n--) {
    uint32_t u = rng_tag(seed, 3u, 0u, (uint32_t)n, 3u);
    int64_t k = (int64_t)(((uint64_t)u * (uint64_t)n) >> 32);
    int t = perm[n - 1]; perm[n - 1] = perm[k]; perm[k] = t;
  }
  for (int64_t r = 0; r < nn; r++)
    for (int64_t c = 0; c < L; c++) sh[r * L + c] = all[r * L + perm[c]];
  /* replay substitutions branch by branch, left then right */
  for (int v = 0; v < N - 1; v++) {
    for (int side = 0; side < 2; side++) {
      int d = side ? right[v] : left[v];
      int64_t pa = N + v, ch = (d <= 0) ? -d : N + d;
      int nsub[25] = {0};
      for (int64_t c = 0; c < L; c++) {
        uint8_t a = all[pa * L + c], b = all[ch * L + c];
        if (a != b && a <= 4 && b <= 4) nsub[a * 5 + b]++;
        sh[ch * L + c] = sh[pa * L + c];
      }
      for (int oldc = 0; oldc < 5; oldc++) {
        int nst = 0;
        for (int nw = 0; nw < 5; nw++) nst += nsub[oldc * 5 + nw];
        if (!nst) continue;
        int ncol = 0;
        for (int64_t c = 0; c < L; c++) if (sh[pa * L + c] == oldc) cols[ncol++] = (int)c;
        if (!ncol) continue;
        for (int k = 0; k < ncol; k++) cperm[k] = k;
        for (int n = ncol; n > 1; n--) {
          uint32_t u = rng_tag(seed, 4u, (uint32_t)(ch * 5 + oldc), (uint32_t)n, 4u + draw);
          int k = (int)(((uint64_t)u * (uint64_t)n) >> 32);
          int t = cperm[n - 1]; cperm[n - 1] = cperm[k]; cperm[k] = t;
        }
        int idx = ncol - 1;
        for (int nw = 0; nw < 5; nw++) {
          int s = nsub[oldc * 5 + nw];
          while (s > 0 && idx >= 0) { sh[ch * L + cols[cperm[idx]]] = (uint8_t)nw; idx--; s--; }
        }
      }
    }
  }
  if (all_out) memcpy(all_out, all, (size_t)(nn * L));
  memcpy(sh_out, sh, (size_t)(nn * L));
  free(all); free(sh); free(S); free(parent); free(perm); free(cols); free(cperm);
  return REF_OK;
}


/* Tree_Substitutions (src/msatree.c:1423-1555) on the Fitch reconstruction of
 * ref_fitch_null_rows (leaves 0..N-1, internal node v at N+v): per column the branches whose
 * child differs from the parent (both canonical unless includegaps), per pair i < j the
 * branches differing at both columns (ndouble) or at either (njoin), [i * L + j]. */
int ref_tree_substitutions(const uint8_t *ax, int64_t N, int64_t L, const int *left, const int *right, uint64_t seed,
                           int includegaps, int32_t *nsubs, int32_t *ndouble, int32_t *njoin)
{
  int64_t nn = 2 * N - 1;
  uint8_t *all = malloc((size_t)(nn * L)), *sh = malloc((size_t)(nn * L));
  if (!all || !sh) { free(all); free(sh); return REF_FAIL; }
  ref_fitch_null_rows(ax, N, L, left, right, seed, all, sh);
  free(sh);
  const int K = 4;
  if (nsubs) {
    for (int64_t i = 0; i < L; i++) {
      nsubs[i] = 0;
      for (int v = 0; v < N - 1; v++) {
        const uint8_t *a  = all + (N + v) * L;
        const uint8_t *al = (left[v] > 0) ? all + (N + left[v]) * L : all + (int64_t)(-left[v]) * L;
        const uint8_t *ar = (right[v] > 0) ? all + (N + right[v]) * L : all + (int64_t)(-right[v]) * L;
        if (includegaps) {
          if (al[i] != a[i]) nsubs[i]++;
          if (ar[i] != a[i]) nsubs[i]++;
        } else if (a[i] < K) {
          if (al[i] < K && al[i] != a[i]) nsubs[i]++;
          if (ar[i] < K && ar[i] != a[i]) nsubs[i]++;
        }
      }
    }
  }
  for (int64_t i = 0; i < L; i++)
    for (int64_t j = 0; j < L; j++) {
      int64_t idx = i * L + j;
      int32_t nd = 0, nj = 0;
      if (i < j)
        for (int v = 0; v < N - 1; v++) {
          const uint8_t *a  = all + (N + v) * L;
          const uint8_t *al = (left[v] > 0) ? all + (N + left[v]) * L : all + (int64_t)(-left[v]) * L;
          const uint8_t *ar = (right[v] > 0) ? all + (N + right[v]) * L : all + (int64_t)(-right[v]) * L;
          if (includegaps) {
            if (al[i] != a[i] && al[j] != a[j]) nd++;
            if (ar[i] != a[i] && ar[j] != a[j]) nd++;
            if (al[i] != a[i] || al[j] != a[j]) nj++;
            if (ar[i] != a[i] || ar[j] != a[j]) nj++;
          } else if (a[i] < K && a[j] < K) {
            if (al[i] < K && al[j] < K && al[i] != a[i] && al[j] != a[j]) nd++;
            if (ar[i] < K && ar[j] < K && ar[i] != a[i] && ar[j] != a[j]) nd++;
            if (al[i] < K && al[j] < K && (al[i] != a[i] || al[j] != a[j])) nj++;
            if (ar[i] < K && ar[j] < K && (ar[i] != a[i] || ar[j] != a[j])) nj++;
          }
        }
      if (ndouble) ndouble[idx] = nd;
      if (njoin) njoin[idx] = nj;
    }
  free(all);
  return REF_OK;
}

/* ---- indels: cov_GenerateAlignmentAddGaps (cov_simulate.c:161-190) ----
 * e1_model_transitions_AF (e1_model.c:784-868) in e2_GLOBAL mode, then e1_model_RenormStateE
 * (:214-270): gamma_X = 1 - exp(-muA_X t) (gammaMDI_func, :1068-1080); beta from
 * e1_model_LI_BetaFunc (:634-690).  rates = {rI, rM, rD, muAM, muAD, muAI, ldI, muI} (AIF:
 * muAD = muAI = muAM, rM = rD = 0, e1_rate.c:431-519).  t16 in the E1_MODEL order: rows B, S, D, I
 * of {S, D, I, E}.  L: the parent's length, used only when beta = 1. */
int ref_e1_transitions(const double *rates, double time, int64_t L, double *t)
{
  const double rI = rates[0], rM = rates[1], rD = rates[2], muAM = rates[3], muAD = rates[4], muAI = rates[5];
  const double ldI = rates[6], muI = rates[7];
  const double gM = 1.0 - exp(-muAM * time), gD = 1.0 - exp(-muAD * time), gI = 1.0 - exp(-muAI * time);
  double beta = 0.0;
  if (time != 0.0 && ldI != 0.0) {
    const double a = ldI - muI, alpha = fabs(a);
    double num, den;
    if (fabs(muI - ldI) < 1e-20) { num = muI * time; den = 1.0 + num; }
    else {
      const double expoI = -time * alpha;
      if (a < 0.) {
        const double A = (ldI > 0.) ? exp(expoI + log(ldI)) : 0.0;
        num = ldI - A; den = muI - A;
      } else {
        const double Al = (ldI > 0.) ? exp(expoI + log(ldI)) : 0.0, Am = (muI > 0.) ? exp(expoI + log(muI)) : 0.0;
        num = ldI - Al; den = ldI - Am;
      }
    }
    beta = (fabs(den) > 0.) ? num / den : 1e-8;
    if (isnan(beta)) return REF_FAIL;
    if (beta > 1.) { if (beta < 1.0 + 1e-3) beta = 1.0; else return REF_FAIL; }
    if (beta < 0.) { if (beta > -1e-4) beta = 0.0; else return REF_FAIL; }
  }
  const double move = 1.0 - (double)L / ((double)L + 1.0);
  t[0] = (1.0 - beta) * (1.0 - gM);                     /* BS */
  t[4] = (1.0 - rM) * (1.0 - beta) * (1.0 - gM) + rM;   /* SS */
  t[8] = (1.0 - rD) * (1.0 - beta) * (1.0 - gD);        /* DS */
  t[12] = (1.0 - rI) * (1.0 - beta) * (1.0 - gI);       /* IS */
  t[1] = (1.0 - beta) * gM;                             /* BD */
  t[5] = (1.0 - rM) * (1.0 - beta) * gM;                /* SD */
  t[9] = (1.0 - rD) * (1.0 - beta) * gD + rD;           /* DD */
  t[13] = (1.0 - rI) * (1.0 - beta) * gI;               /* ID */
  t[2] = (beta == 1.0) ? 1.0 - move : beta;             /* BI */
  t[6] = ((1.0 - rM) * beta == 1.0) ? 1.0 - move : (1.0 - rM) * beta;
  t[10] = ((1.0 - rD) * beta == 1.0) ? 1.0 - move : (1.0 - rD) * beta;
  t[14] = ((1.0 - rI) * beta == 1.0) ? 1.0 - move : (1.0 - rI) * beta + rI;
  for (int r = 0; r < 4; r++) {                          /* RenormStateE: X -> E is 0 */
    double sum = t[4 * r] + t[4 * r + 1] + t[4 * r + 2];
    sum = (sum > 0.) ? 1.0 / sum : 1.0;
    for (int k = 0; k < 3; k++) t[4 * r + k] *= sum;
    t[4 * r + 3] = 0.0;
  }
  return REF_OK;
}

static uint32_t th32(double x) { return (x >= 1.0) ? 0xFFFFFFFFu : (uint32_t)(x * 4294967296.0); }

/* The indel pass on the leaves (cov_evolve_indels_tree, :326-361; the branch step, :495-563;
 * cov_emit_indels, :634-722; cov_insert, :803-859): branches in tree order (v = 0..N-2, left then
 * right child), each walking the e1 HMM over the current alignment (all its columns, those earlier
 * branches inserted included): from B, every step draws one uniform (key: child row, step, tag 3)
 * and moves to S, D or I by the current state's row of the branch's transitions (esl_rnd_FChoose
 * over the three); S and D consume a column (D gaps the child's cell), I inserts a column right
 * after the current one (gap in every other row; the child's residue drawn from insf with key
 * (child row, insertion ordinal, tag 4)); the walk ends when a step consumes column L + 1.  Only
 * leaf rows are kept: internal rows are never output, and their branches' insertions add gap
 * columns to every leaf.  ct follows the columns (inserted columns unpaired).  leaves_in N x L0,
 * leaves_out N x cap; returns REF_FAIL when the alignment would outgrow cap. */
int ref_add_indels(int64_t N, int64_t L0, const int *left, const int *right, const double *ld, const double *rd,
                   const double *rates, const double *insf, uint64_t seed, uint32_t stream, const uint8_t *leaves_in,
                   const int *ct_in, int64_t cap, uint8_t *leaves_out, int *ct_out, int64_t *alen)
{
  if (L0 > cap) return REF_FAIL;
  int64_t len = L0;
  for (int64_t l = 0; l < N; l++) memcpy(leaves_out + l * cap, leaves_in + l * L0, (size_t)L0);
  if (ct_out) for (int64_t c = 0; c < L0; c++) ct_out[c] = ct_in ? ct_in[c] : -1;
  uint32_t ith[3];
  double cum = 0., tot = insf[0] + insf[1] + insf[2] + insf[3];
  for (int b = 0; b < 3; b++) { cum += insf[b]; ith[b] = th32(cum / tot); }
  for (int64_t v = 0; v < N - 1; v++)
    for (int side = 0; side < 2; side++) {
      const int d = side ? right[v] : left[v];
      const int64_t row = (d > 0) ? d : (N - 1 - d), leaf = (d > 0) ? -1 : -d;
      double t16[16];
      if (ref_e1_transitions(rates, side ? rd[v] : ld[v], len, t16) != REF_OK) return REF_FAIL;
      uint32_t th[4][2];
      for (int r = 0; r < 4; r++) {
        const double s = t16[4 * r] + t16[4 * r + 1] + t16[4 * r + 2];
        th[r][0] = th32(t16[4 * r] / s); th[r][1] = th32((t16[4 * r] + t16[4 * r + 1]) / s);
      }
      const int64_t L = len;
      int64_t i = 0, pos = 0, step = 0, nins = 0;
      int st = 0;                                          /* B */
      for (;;) {
        const uint32_t u = rng_tag(seed, stream, (uint32_t)row, (uint32_t)step++, 3u);
        st = (u < th[st][0]) ? 1 : (u < th[st][1]) ? 2 : 3;  /* S, D, I */
        if (st == 1 || st == 2) { i++; pos++; }
        if (i == L + 1) break;
        if (st == 2 && leaf >= 0) leaves_out[leaf * cap + pos - 1] = 4;
        if (st == 3) {                                     /* new column at index pos (after 1-based pos) */
          if (len + 1 > cap) return REF_FAIL;
          for (int64_t l = 0; l < N; l++) {
            uint8_t *r8 = leaves_out + l * cap;
            memmove(r8 + pos + 1, r8 + pos, (size_t)(len - pos));
            r8[pos] = 4;
          }
          if (leaf >= 0)
            leaves_out[leaf * cap + pos] = (uint8_t)icdf(rng_tag(seed, stream, (uint32_t)row, (uint32_t)nins, 4u), ith, 4);
          if (ct_out) {
            memmove(ct_out + pos + 1, ct_out + pos, sizeof(int) * (size_t)(len - pos));
            ct_out[pos] = -1;
            for (int64_t c = 0; c <= len; c++) if (ct_out[c] >= pos) ct_out[c]++;
          }
          len++; nins++; pos++;
        }
      }
    }
  *alen = len;
  return REF_OK;
}
