/* ref_rank.c -- oracle: score histograms, null merge, tail fits, E-values.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * R-scape's own code (src/covariation.c, src/R-scape.c) is restated loop for
 * loop.  The histogram object and the fits belong to Easel, which is not
 * vendored in the reference tree (src/Makefile:58-60; version not pinned).
 * They are restated from Easel's published semantics:
 *   - bins are left-open / right-closed: b = ceil((x - bmin)/w - 1), bin b spans
 *     (bmin + b w, bmin + (b+1) w];  lower / upper bound = bmin + b w / bmin + (b+1) w;
 *   - the histogram grows below / above by 2x over-allocation;
 *   - SetTailByMass: walk down from imax until the kept count >= pmass * n,
 *     censor virtually at the lower bound phi of that bin;
 *   - binned exponential MLE, closed form lambda = (ln Sb - ln Sa)/w with
 *     Sa = sum n_b (a_b - mu), Sb = sum n_b (b_b - mu), mu = phi;
 *   - binned Gamma MLE: PARITY UNPINNED (method restated: ML on bin midpoints
 *     above mu = phi, tau solved from ln tau - digamma(tau) = ln mean - mean ln).
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "refcpu.h"

ref_hist *ref_hist_create(double bmin, double bmax, double w)
{
  ref_hist *h = calloc(1, sizeof(ref_hist));
  if (!h) return NULL;
  h->bmin = bmin; h->bmax = bmax; h->w = w;
  h->nb   = (int)((bmax - bmin) / w);
  h->obs  = calloc((size_t)(h->nb > 0 ? h->nb : 1), sizeof(uint64_t));
  h->imin = h->nb; h->imax = -1; h->cmin = h->imin;
  h->xmin = DBL_MAX; h->xmax = -DBL_MAX; h->phi = 0.;
  return h;
}

void ref_hist_destroy(ref_hist *h) { if (h) { free(h->obs); free(h); } }

int ref_hist_score2bin(const ref_hist *h, double x, int *ret_b)
{
  if (!isfinite(x)) return REF_FAIL;
  double b = ceil(((x - h->bmin) / h->w) - 1.);
  if (b < (double)INT_MIN || b > (double)INT_MAX) return REF_FAIL;
  *ret_b = (int)b;
  return REF_OK;
}

double ref_hist_lbound(const ref_hist *h, int b) { return h->w * b + h->bmin; }
double ref_hist_ubound(const ref_hist *h, int b) { return h->w * (b + 1) + h->bmin; }

int ref_hist_add(ref_hist *h, double x)
{
  int b;
  if (ref_hist_score2bin(h, x, &b) != REF_OK) return REF_FAIL;
  if (b < 0) {
    int nnew = -b * 2;
    uint64_t *o = calloc((size_t)(nnew + h->nb), sizeof(uint64_t));
    memcpy(o + nnew, h->obs, sizeof(uint64_t) * (size_t)h->nb);
    free(h->obs); h->obs = o;
    h->nb += nnew; b += nnew; h->bmin -= nnew * h->w;
    h->imin += nnew; h->cmin += nnew;
    if (h->imax > -1) h->imax += nnew;
  } else if (b >= h->nb) {
    int nnew = (b - h->nb + 1) * 2;
    uint64_t *o = calloc((size_t)(nnew + h->nb), sizeof(uint64_t));
    memcpy(o, h->obs, sizeof(uint64_t) * (size_t)h->nb);
    free(h->obs); h->obs = o;
    if (h->imin == h->nb) { h->imin += nnew; h->cmin += nnew; }
    h->bmax += nnew * h->w;
    h->nb   += nnew;
  }
  h->obs[b]++;
  h->n++; h->Nc++; h->No++;
  if (b > h->imax) h->imax = b;
  if (b < h->imin) { h->imin = b; h->cmin = b; }
  if (x > h->xmax) h->xmax = x;
  if (x < h->xmin) h->xmin = x;
  return REF_OK;
}

int ref_hist_set_tail_by_mass(ref_hist *h, double pmass, double *ret_newmass)
{
  int b;
  uint64_t sum = 0;
  for (b = h->imax; b >= h->imin; b--) {
    sum += h->obs[b];
    if ((double)sum >= pmass * (double)h->n) break;
  }
  h->phi = ref_hist_lbound(h, b);
  h->cmin = b;
  h->Nc = h->n;
  h->No = sum;
  h->tailfit = 1;
  if (ret_newmass) *ret_newmass = (double)sum / (double)h->n;
  return REF_OK;
}

int ref_exp_fit_binned(const ref_hist *h, double *ret_mu, double *ret_lambda)
{
  double mu; int i;
  if (h->tailfit) { mu = h->phi; i = h->cmin; }
  else            { mu = h->xmin; i = h->imin; }
  double sa = 0., sb = 0.;
  for (; i <= h->imax; i++) {
    double ai = ref_hist_lbound(h, i), bi = ref_hist_ubound(h, i);
    sa += (double)h->obs[i] * (ai - mu);
    sb += (double)h->obs[i] * (bi - mu);
  }
  *ret_mu = mu;
  *ret_lambda = 1. / h->w * (log(sb) - log(sa));
  return REF_OK;
}

static double digamma(double x)
{
  double r = 0.;
  while (x < 6.) { r -= 1. / x; x += 1.; }
  double f = 1. / (x * x);
  return r + log(x) - 0.5 / x - f * (1. / 12 - f * (1. / 120 - f * (1. / 252 - f * (1. / 240 - f / 132))));
}

int ref_gam_fit_binned(const ref_hist *h, double *ret_mu, double *ret_lambda, double *ret_tau)
{
  double mu; int i0;
  if (h->tailfit) { mu = h->phi; i0 = h->cmin; }
  else            { mu = h->xmin; i0 = h->imin; }
  double n = 0., mean = 0., lsum = 0.;
  for (int i = i0; i <= h->imax; i++) {
    if (!h->obs[i]) continue;
    double x = 0.5 * (ref_hist_lbound(h, i) + ref_hist_ubound(h, i)) - mu;
    if (x <= 0.) continue;
    n += (double)h->obs[i]; mean += (double)h->obs[i] * x; lsum += (double)h->obs[i] * log(x);
  }
  *ret_mu = mu;
  if (n <= 0.) { *ret_lambda = INFINITY; *ret_tau = 0.; return REF_OK; }
  mean /= n; lsum /= n;
  double s = log(mean) - lsum;              /* >= 0 by Jensen */
  if (!(s > 1e-12)) { *ret_lambda = INFINITY; *ret_tau = INFINITY; return REF_OK; }
  double lo = 1e-6, hi = 1e6;               /* f(tau) = ln tau - digamma(tau) - s is decreasing */
  for (int it = 0; it < 200; it++) {
    double mid = sqrt(lo * hi);
    double f = log(mid) - digamma(mid) - s;
    if (f > 0) lo = mid; else hi = mid;
    if (hi / lo < 1. + 1e-14) break;
  }
  double tau = sqrt(lo * hi);
  *ret_tau = tau;
  *ret_lambda = tau / mean;
  return REF_OK;
}

double ref_exp_surv(double x, double mu, double lambda)
{
  if (x < mu) return 1.0;
  return exp(-lambda * (x - mu));
}

/* Regularized upper incomplete gamma Q(a,x): series for x < a+1, continued fraction otherwise
 * (Press et al., Numerical Recipes, 2nd ed., section 6.2). */
static double gammq(double a, double x)
{
  if (x <= 0.) return 1.0;
  double gln = lgamma(a);
  if (x < a + 1.) {
    double ap = a, sum = 1. / a, del = sum;
    for (int n = 0; n < 1000; n++) { ap += 1.; del *= x / ap; sum += del; if (fabs(del) < fabs(sum) * 1e-16) break; }
    return 1. - sum * exp(-x + a * log(x) - gln);
  }
  double b = x + 1. - a, c = 1. / DBL_MIN, d = 1. / b, hh = d;
  for (int i = 1; i < 1000; i++) {
    double an = -i * (i - a);
    b += 2.;
    d = an * d + b; if (fabs(d) < DBL_MIN) d = DBL_MIN;
    c = b + an / c; if (fabs(c) < DBL_MIN) c = DBL_MIN;
    d = 1. / d;
    double del = d * c; hh *= del;
    if (fabs(del - 1.) < 1e-16) break;
  }
  return exp(-x + a * log(x) - gln) * hh;
}

double ref_gam_surv(double x, double mu, double lambda, double tau)
{
  if (x < mu) return 1.0;
  return gammq(tau, lambda * (x - mu));
}

/* covariation.c:1676-1699: survfit[b] = pmass * S(UBound(b)) for b in [cmin, 2 nb). */
double *ref_survfit(const ref_hist *h, double pmass, int gamma, const double *ep)
{
  double *sf = calloc((size_t)(2 * h->nb), sizeof(double));
  if (!sf) return NULL;
  for (int b = h->cmin; b < 2 * h->nb; b++) {
    double bi = ref_hist_ubound(h, b);
    sf[b] = pmass * (gamma ? ref_gam_surv(bi, ep[0], ep[1], ep[2]) : ref_exp_surv(bi, ep[0], ep[1]));
  }
  return sf;
}

/* covariation.c:2483-2506 */
double ref_hist_pmass(const ref_hist *h, double target_pmass, double target_fracfit)
{
  int tp = 0, nfit = 0;
  uint64_t c = 0;
  double pmass = NAN;
  for (int i = h->imax; i >= h->imin; i--) if (h->obs[i] > 0) tp++;
  for (int i = h->imax; i >= h->imin; i--) {
    c += h->obs[i];
    if (h->obs[i] > 0) {
      nfit++;
      pmass = (double)c / (double)h->Nc;
      if ((double)nfit / (double)tp >= target_fracfit || pmass >= target_pmass) break;
    }
  }
  return pmass;
}

/* covariation.c:2369-2401 (note survfit[icov+1]; the reference exits on an
 * impossible bin, here that returns NaN). */
double ref_cov2evalue(double cov, int64_t Nc, const ref_hist *h, const double *survfit)
{
  double eval = INFINITY;
  int icov;
  uint64_t c = 0;
  ref_hist_score2bin(h, cov, &icov);
  if (survfit && icov >= 2 * h->nb - 1)     eval = survfit[2 * h->nb - 1] * (double)Nc;
  else if (survfit && cov >= h->phi)        eval = survfit[icov + 1] * (double)Nc;
  else {
    if (cov >= h->xmax) return (double)Nc / (double)h->Nc;
    if (icov <= h->imax) {
      if (icov < h->imin) icov = h->imin;
      if (icov >= h->imax - 1) eval = (double)Nc / (double)h->Nc;
      else {
        for (int i = h->imax; i >= icov; i--) c += h->obs[i];
        eval = (double)c * (double)Nc / (double)h->Nc;
      }
    } else return NAN;
  }
  return eval;
}

/* covariation.c:2404-2436 */
double ref_evalue2cov(double eval_thresh, int64_t Nc, const ref_hist *h, const double *survfit)
{
  double cov = -INFINITY, eval;
  int min_nobs = 10, i, b = h->cmin - 1;
  uint64_t c = 0;
  if (h->No >= (uint64_t)min_nobs && survfit) {
    for (b = 2 * h->nb - 1; b >= h->cmin; b--) {
      eval = survfit[b] * (double)Nc;
      if (eval >= eval_thresh) break;
    }
    cov = ref_hist_lbound(h, b);
  }
  if (b == h->cmin - 1) {
    for (i = h->imax; i >= h->imin; i--) {
      c += h->obs[i];
      eval = (double)c * (double)Nc / (double)h->Nc;
      if (eval >= eval_thresh) break;
    }
    cov = ref_hist_ubound(h, i);
  }
  return cov;
}

/* covariation.c:414-432: bmax = maxCOV + 5w (bumped while |bmax-bmin| < tol),
 * histogram of max(COV_ij, bmin + w) over i<j. */
ref_hist *ref_rank_hist(int64_t L, const double *COV, double maxCOV, double bmin, double w, double tol)
{
  double bmax = maxCOV + 5 * w;
  while (fabs(bmax - bmin) < tol) bmax += w;
  ref_hist *h = ref_hist_create(bmin, bmax, w);
  for (int64_t i = 0; i < L - 1; i++)
    for (int64_t j = i + 1; j < L; j++) {
      double v = COV[i * L + j];
      double add = (v > bmin + w) ? v : bmin + w;
      ref_hist_add(h, add);
    }
  return h;
}

/* covariation.c:2333-2364 */
static int bin2bin(int b, const ref_hist *h, const ref_hist *nw, int *ret_newb)
{
  double x = ref_hist_lbound(h, b);
  *ret_newb = -1;
  if (!isfinite(x)) return REF_FAIL;
  x = round((x - nw->bmin) / nw->w);
  int nb = (int)x;
  if (nb > nw->nb) { *ret_newb = -1; return REF_FAIL; }
  *ret_newb = nb;
  return REF_OK;
}

/* covariation.c:682-736 (GrowRankList, ha only) */
static ref_hist *grow(ref_hist *old, double bmax, double bmin)
{
  double new_bmin = old->bmin;
  if (bmin < old->bmin) new_bmin -= fabs(bmin) * 2. * old->w;
  ref_hist *nw = ref_hist_create(new_bmin, (bmax > old->bmax) ? bmax : old->bmax, old->w);
  nw->n = old->n; nw->xmin = old->xmin; nw->xmax = old->xmax;
  nw->imin = old->imin; nw->imax = old->imax; nw->Nc = old->Nc; nw->No = old->No;
  for (int b = old->imin; b <= old->imax; b++) {
    int nb;
    bin2bin(b, old, nw, &nb);
    if (nb >= 0 && nb < nw->nb) nw->obs[nb] = old->obs[b];
  }
  ref_hist_destroy(old);
  return nw;
}

/* R-scape.c:1564-1612 (null_add2cumranklist, ha only) */
int ref_null_add(ref_hist **pcum, const ref_hist *h)
{
  ref_hist *cum = *pcum;
  if (!cum) {
    cum = ref_hist_create(h->bmin, h->bmax, h->w);
    cum->n = h->n; cum->xmin = h->xmin; cum->xmax = h->xmax;
    cum->imin = h->imin; cum->imax = h->imax;
  } else {
    cum = grow(cum, h->bmax, h->bmin);
    cum->n += h->n;
    if (h->xmin < cum->xmin) cum->xmin = h->xmin;
    if (h->xmax > cum->xmax) cum->xmax = h->xmax;
    if (h->imin < cum->imin) cum->imin = h->imin;
    if (h->imax > cum->imax) cum->imax = h->imax;
  }
  for (int b = h->imin; b <= h->imax; b++) {
    int cb;
    bin2bin(b, h, cum, &cb);
    if (cb >= 0 && cb < cum->nb) { cum->obs[cb] += h->obs[b]; cum->Nc += h->obs[b]; cum->No += h->obs[b]; }
  }
  *pcum = cum;
  return REF_OK;
}

/* R-scape.c:1353-1360: w = min(w_old, (maxCOV - max(bmin, minCOV))/hpts); w < tol -> 0. */
double ref_width(double maxCOV, double minCOV, double bmin, double w_old, int hpts, double tol)
{
  double lo = (bmin > minCOV) ? bmin : minCOV;
  double w_new = (maxCOV - lo) / (double)hpts;
  double w = (w_old < w_new) ? w_old : w_new;
  if (w < tol) w = 0.0;
  return w;
}

/* covariation.c:459-488 + :1914-1975: pmass, SetTailByMass, fit, survfit.
 * ret_fit = {mu, lambda, tau, newmass, pmass}. */
int ref_null_fit(ref_hist *null, double pmass_target, double fracfit, int doexpfit,
                 double **ret_survfit, double *ret_fit)
{
  double pmass = ref_hist_pmass(null, pmass_target, fracfit);
  if (isnan(pmass)) return REF_FAIL;
  double newmass, ep[3] = {0., 0., 0.};
  ref_hist_set_tail_by_mass(null, pmass, &newmass);
  if (doexpfit) ref_exp_fit_binned(null, &ep[0], &ep[1]);
  else          ref_gam_fit_binned(null, &ep[0], &ep[1], &ep[2]);
  *ret_survfit = NULL;
  if (!isinf(ep[1])) *ret_survfit = ref_survfit(null, newmass, !doexpfit, ep);
  ret_fit[0] = ep[0]; ret_fit[1] = ep[1]; ret_fit[2] = ep[2]; ret_fit[3] = newmass; ret_fit[4] = pmass;
  return REF_OK;
}

/* covariation.c:846-856: p = cov2evalue(cov, 1, ...); E = p * Nc for every i<j. */
int ref_evalues(int64_t L, const double *COV, int64_t Nc, const ref_hist *null,
                const double *survfit, double *Eval)
{
  for (int64_t c = 0; c < L * L; c++) Eval[c] = INFINITY;
  for (int64_t i = 0; i < L - 1; i++)
    for (int64_t j = i + 1; j < L; j++) {
      double p = ref_cov2evalue(COV[i * L + j], 1, null, survfit);
      Eval[i * L + j] = Eval[j * L + i] = p * (double)Nc;
    }
  return REF_OK;
}


/* cov_WriteNullHistogram (covariation.c:1847-1868): one "%f %llu" line per bin, the bin's lower
 * bound bmin + i w and its count. */
int ref_write_null(const char *path, const ref_hist *h)
{
  FILE *fp = fopen(path, "w");
  if (!fp) return REF_FAIL;
  for (int i = 0; i < h->nb; i++) fprintf(fp, "%f %llu\n", h->bmin + i * h->w, (unsigned long long)h->obs[i]);
  fclose(fp);
  return REF_OK;
}

/* cov_ReadNullHistogram (covariation.c:1717-1844).  Pass 1 over every line's first token: w is
 * the smallest |score - previous score| (the first line compares with -inf), bmin / bmax the
 * score extremes (bmax grown by w while within 1e-2 of bmin).  Pass 2: a line with exactly two
 * tokens sets the bin Score2Bin(score) -- a written lower bound lands one bin below the bin it
 * was written from -- to its count (atoi), growing the histogram below / above as Easel does. */
ref_hist *ref_read_null(const char *path)
{
  FILE *fp = fopen(path, "r");
  if (!fp) return NULL;
  char line[1024];
  double w = INFINITY, sc = -INFINITY, mx = -INFINITY, mn = INFINITY;
  while (fgets(line, sizeof line, fp)) {
    if (line[0] == '#') continue;
    char *tok = strtok(line, " \t\n");
    if (!tok) continue;
    double prv = sc;
    sc = atof(tok);
    if (fabs(sc - prv) < w) w = fabs(sc - prv);
    if (sc > mx) mx = sc;
    if (sc < mn) mn = sc;
  }
  double bmax = mx, bmin = mn;
  while (fabs(bmax - bmin) < 1e-2) bmax += w;
  ref_hist *h = ref_hist_create(bmin, bmax, w);
  rewind(fp);
  while (fgets(line, sizeof line, fp)) {
    if (line[0] == '#') continue;
    int t = 0;
    double x = 0.;
    uint64_t counts = 0;
    for (char *tok = strtok(line, " \t\n"); tok; tok = strtok(NULL, " \t\n"), t++) {
      if (t == 0) x = atof(tok);
      else        counts = (uint64_t)atoi(tok);
    }
    if (t != 2) continue;
    int b;
    if (ref_hist_score2bin(h, x, &b) != REF_OK) { ref_hist_destroy(h); fclose(fp); return NULL; }
    if (b < 0) {
      int nnew = -b * 2;
      uint64_t *o = calloc((size_t)(nnew + h->nb), sizeof(uint64_t));
      memcpy(o + nnew, h->obs, sizeof(uint64_t) * (size_t)h->nb);
      free(h->obs); h->obs = o;
      h->nb += nnew; b += nnew; h->bmin -= nnew * h->w;
      h->imin += nnew; h->cmin += nnew;
      if (h->imax > -1) h->imax += nnew;
    } else if (b >= h->nb) {
      int nnew = (b - h->nb + 1) * 2;
      uint64_t *o = calloc((size_t)(nnew + h->nb), sizeof(uint64_t));
      memcpy(o, h->obs, sizeof(uint64_t) * (size_t)h->nb);
      free(h->obs); h->obs = o;
      if (h->imin == h->nb) { h->imin += nnew; h->cmin += nnew; }
      h->bmax += nnew * h->w;
      h->nb   += nnew;
    }
    if (b < h->imin) { h->imin = b; h->cmin = b; }
    if (b > h->imax) h->imax = b;
    if (x < h->xmin) h->xmin = x;
    if (x > h->xmax) h->xmax = x;
    h->obs[b] = counts;
    h->n += counts; h->Nc += counts; h->No += counts;
  }
  fclose(fp);
  return h;
}
