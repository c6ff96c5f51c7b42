// rsg_rank.hip -- score histograms, null histogram merge, tail fits and E-values.
//
// Reference (R-scape v2.6.11): cov_SignificantPairs_Ranking src/covariation.c:325-530,
// cov_CreateRankList :640-665, cov_GrowRankList :682-736, null histogram I/O :1717-1868,
// cov_NullFitExponential :1914-1941, cov_NullFitGamma :1943-1975, SetSurvFitTail :1676-1699,
// cov2evalue / evalue2cov :2369-2436, cov_histogram_pmass :2483-2506,
// cov_ranklist_Bin2Bin :2333-2364; calculate_width_histo src/R-scape.c:1280-1371 and
// null_add2cumranklist :1564-1612.
//
// The histogram follows Easel's esl_histogram (not vendored in the reference tree):
// nb = (int)((bmax - bmin)/w) bins, bin b = ceil((x - bmin)/w - 1) covers
// (bmin + b w, bmin + (b+1) w], growth by 2x over-allocation below or above, and
// SetTailByMass censors virtually at the lower bound of the bin where the kept mass
// first reaches pmass * n.  The exponential tail fit is the binned ML closed form
// lambda = (ln Sb - ln Sa)/w; the Gamma fit is ML on bin midpoints above phi.
//
// Device work (FP64): the histogram of a COV matrix, the pmass / tail / fit / survival
// computation (one work group, parallel scans and reductions), and the E-value of every
// pair.  Histogram merging and the .null text format are host code.
#include "rsg_device.h"
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <climits>
#include <cfloat>
#include <algorithm>

// Histogram geometry, bins and E-values are rounded exactly as the host computes them.
#pragma clang fp contract(off)

// ---------------------------------------------------------------- host histogram object
extern "C" rsg_hist *rsg_hist_create(double bmin, double bmax, double w)
{
  rsg_hist *h = (rsg_hist *)calloc(1, sizeof(rsg_hist));
  if (!h) return nullptr;
  h->bmin = bmin; h->bmax = bmax; h->w = w;
  h->nb = (int32_t)((bmax - bmin) / w);
  h->obs = (uint64_t *)calloc((size_t)std::max(h->nb, 1), sizeof(uint64_t));
  h->imin = h->nb; h->imax = -1; h->cmin = h->imin;
  h->xmin = DBL_MAX; h->xmax = -DBL_MAX; h->phi = 0.;
  return h;
}

extern "C" void rsg_hist_destroy(rsg_hist *h) { if (h) { free(h->obs); free(h); } }

static int score2bin(const rsg_hist *h, double x, int *b)
{
  if (!std::isfinite(x)) return RSG_FAIL;
  double v = ceil(((x - h->bmin) / h->w) - 1.);
  if (v < (double)INT_MIN || v > (double)INT_MAX) return RSG_FAIL;
  *b = (int)v;
  return RSG_OK;
}
static double lbound(const rsg_hist *h, int b) { return h->w * b + h->bmin; }
static double ubound(const rsg_hist *h, int b) { return h->w * (b + 1) + h->bmin; }

// make bin b exist (growing below or above as Easel does); returns the shifted bin index
static int grow_to(rsg_hist *h, int b)
{
  if (b < 0) {
    int nnew = -b * 2;
    uint64_t *o = (uint64_t *)calloc((size_t)(nnew + h->nb), sizeof(uint64_t));
    memcpy(o + nnew, h->obs, sizeof(uint64_t) * (size_t)h->nb);
    free(h->obs); h->obs = o;
    h->nb += nnew; b += nnew; h->bmin -= nnew * h->w;
    h->imin += nnew; h->cmin += nnew;
    if (h->imax > -1) h->imax += nnew;
  } else if (b >= h->nb) {
    int nnew = (b - h->nb + 1) * 2;
    uint64_t *o = (uint64_t *)calloc((size_t)(nnew + h->nb), sizeof(uint64_t));
    memcpy(o, h->obs, sizeof(uint64_t) * (size_t)h->nb);
    free(h->obs); h->obs = o;
    if (h->imin == h->nb) { h->imin += nnew; h->cmin += nnew; }
    h->bmax += nnew * h->w;
    h->nb += nnew;
  }
  return b;
}

extern "C" double rsg_width(double maxCOV, double minCOV, double bmin, double w_old, int hpts, double tol)
{
  double lo = std::max(bmin, minCOV);
  double w = std::min(w_old, (maxCOV - lo) / (double)hpts);
  return (w < tol) ? 0.0 : w;
}

// ---------------------------------------------------------------- device histogram
__global__ void __launch_bounds__(256)
k_hist(const double *__restrict__ cov, int64_t L, double bmin, double w, int nb, int use_lds,
       unsigned long long *__restrict__ obs)
{
  extern __shared__ unsigned sh[];
  if (use_lds) {
    for (int b = threadIdx.x; b < nb; b += blockDim.x) sh[b] = 0;
    __syncthreads();
  }
  const double floor_v = bmin + w;
  const int64_t n = L * L;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx / L, j = idx - i * L;
    if (i >= j) continue;
    double x = cov[idx];
    double add = (x > floor_v) ? x : floor_v;                 // ESL_MAX(COV, bmin + w)
    int b = (int)ceil(((add - bmin) / w) - 1.);
    b = (b < 0) ? 0 : ((b >= nb) ? nb - 1 : b);               // cannot happen: bmax = maxCOV + 5w
    if (use_lds) atomicAdd(&sh[b], 1u);
    else         atomicAdd(&obs[b], 1ull);
  }
  if (use_lds) {
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
      if (sh[b]) atomicAdd(&obs[b], (unsigned long long)sh[b]);
  }
}

extern "C" int rsg_cov_RankHist(rsg_ctx *c, rsg_msa *m, double bmin, double w, double tol, rsg_hist **ret, char *errbuf)
{
  *ret = nullptr;
  if (!m->have_cov) return rsg_fail(errbuf, "no covariation matrix to rank");
  if (!(w > 0.)) return rsg_fail(errbuf, "histogram width must be positive (w = %g)", w);
  double bmax = m->maxCOV + 5 * w;
  while (fabs(bmax - bmin) < tol) bmax += w;
  // bmin above every score leaves no bin (Easel's histogram would grow; the reference never
  // asks for this, its bmin is -10 and bmax = maxCOV + 5w): an error, before any allocation
  if (!(bmax > bmin) || !std::isfinite(bmax))
    return rsg_fail(errbuf, "histogram range is empty (bmin %g, bmax %g)", bmin, bmax);
  rsg_hist *h = rsg_hist_create(bmin, bmax, w);
  if (!h) return rsg_fail(errbuf, "out of memory");
  const int nb = h->nb;
  if (nb < 1) { rsg_hist_destroy(h); return rsg_fail(errbuf, "histogram has no bins (bmin %g, bmax %g, w %g)", bmin, bmax, w); }
  unsigned long long *d_obs = nullptr;
  const int use_lds = (nb <= 16384);          // LDS-privatised bins up to 64 KB per work group
  const int64_t L = m->L;
  int blocks = (int)std::min<int64_t>(2048, (L * L + 255) / 256);
  int t;
  hipError_t e = hipMalloc(&d_obs, sizeof(unsigned long long) * (size_t)nb);
  if (e == hipSuccess) e = hipMemsetAsync(d_obs, 0, sizeof(unsigned long long) * (size_t)nb, c->stream);
  if (e == hipSuccess) {
    rsg_tstart(c, "hist", &t);
    hipLaunchKernelGGL(k_hist, dim3(blocks), dim3(256), use_lds ? sizeof(unsigned) * nb : 0, c->stream,
                       m->d_cov, L, bmin, w, nb, use_lds, d_obs);
    rsg_tstop(c, t);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(h->obs, d_obs, sizeof(uint64_t) * (size_t)nb, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (d_obs) hipFree(d_obs);
  if (e != hipSuccess) { rsg_hist_destroy(h); return rsg_fail(errbuf, "rsg_cov_RankHist: %s", hipGetErrorString(e)); }
  // scalar fields: the added values are max(COV, bmin + w), so their extremes follow
  // from the COV extremes of the upper triangle
  uint64_t npairs = (uint64_t)(L * (L - 1) / 2);
  double fl = bmin + w;
  h->xmin = std::max(m->minCOV, fl);
  h->xmax = std::max(m->maxCOV, fl);
  int bmin_b, bmax_b;
  score2bin(h, h->xmin, &bmin_b); score2bin(h, h->xmax, &bmax_b);
  h->imin = bmin_b; h->imax = bmax_b; h->cmin = bmin_b;
  h->n = h->Nc = h->No = npairs;
  if (npairs == 0) { h->imin = h->nb; h->imax = -1; h->cmin = h->nb; h->xmin = DBL_MAX; h->xmax = -DBL_MAX; }
  *ret = h;
  return RSG_OK;
}

// ---------------------------------------------------------------- null histogram merge
static int bin2bin(const rsg_hist *h, int b, const rsg_hist *nw, int *ret)
{
  double x = lbound(h, b);
  *ret = -1;
  if (!std::isfinite(x)) return RSG_FAIL;
  x = round((x - nw->bmin) / nw->w);
  if (x < (double)INT_MIN || x > (double)INT_MAX) return RSG_FAIL;
  int nb = (int)x;
  if (nb > nw->nb) return RSG_FAIL;
  *ret = nb;
  return RSG_OK;
}

extern "C" int rsg_hist_NullAdd(rsg_hist **pcum, const rsg_hist *h, char *errbuf)
{
  if (!h) return RSG_OK;
  rsg_hist *cum = *pcum;
  if (!cum) {
    cum = rsg_hist_create(h->bmin, h->bmax, h->w);
    if (!cum) return rsg_fail(errbuf, "out of memory");
    cum->n = h->n; cum->xmin = h->xmin; cum->xmax = h->xmax; cum->imin = h->imin; cum->imax = h->imax;
  } else {
    // cov_GrowRankList: a new histogram covering both, bins carried over by Bin2Bin
    double new_bmin = cum->bmin;
    if (h->bmin < cum->bmin) new_bmin -= fabs(h->bmin) * 2. * cum->w;
    rsg_hist *nw = rsg_hist_create(new_bmin, std::max(h->bmax, cum->bmax), cum->w);
    if (!nw) return rsg_fail(errbuf, "out of memory");
    nw->n = cum->n; nw->xmin = cum->xmin; nw->xmax = cum->xmax; nw->imin = cum->imin; nw->imax = cum->imax;
    nw->Nc = cum->Nc; nw->No = cum->No;
    for (int b = cum->imin; b <= cum->imax; b++) {
      int nb2;
      bin2bin(cum, b, nw, &nb2);
      if (nb2 >= 0 && nb2 < nw->nb) nw->obs[nb2] = cum->obs[b];
    }
    rsg_hist_destroy(cum);
    cum = nw;
    cum->n += h->n;
    cum->xmin = std::min(cum->xmin, h->xmin);
    cum->xmax = std::max(cum->xmax, h->xmax);
    cum->imin = std::min(cum->imin, h->imin);
    cum->imax = std::max(cum->imax, h->imax);
  }
  for (int b = h->imin; b <= h->imax; b++) {
    int cb;
    bin2bin(h, b, cum, &cb);
    if (cb >= 0 && cb < cum->nb) { cum->obs[cb] += h->obs[b]; cum->Nc += h->obs[b]; cum->No += h->obs[b]; }
  }
  *pcum = cum;
  return RSG_OK;
}

// ---------------------------------------------------------------- FP64 tail fit on the device
// (body in rsg_device.h: nullfit_body)
__global__ void __launch_bounds__(1024)
k_nullfit(const unsigned long long *__restrict__ obs, int nb, int imin, int imax, double n, double Nc,
          double bmin, double w, double target_pmass, double fracfit, int doexp,
          unsigned long long *__restrict__ S, int *__restrict__ NZ, double *__restrict__ survfit, double *out)
{
  nullfit_body(obs, nb, imin, imax, n, Nc, bmin, w, target_pmass, fracfit, doexp, S, NZ, survfit, out);
}

extern "C" int rsg_hist_NullFit(rsg_ctx *c, rsg_hist *h, double pmass, double fracfit, int doexpfit,
                                double **ret_survfit, double *fit, char *errbuf)
{
  *ret_survfit = nullptr;
  if (!h || h->imax < h->imin) return rsg_fail(errbuf, "empty null histogram");
  const int nb = h->nb;
  unsigned long long *d_obs = nullptr, *d_S = nullptr;
  int *d_NZ = nullptr;
  double *d_sf = nullptr;
  double *d_out = nullptr;
  HIPCHK(hipMalloc(&d_obs, sizeof(unsigned long long) * nb));
  HIPCHK(hipMalloc(&d_S, sizeof(unsigned long long) * nb));
  HIPCHK(hipMalloc(&d_NZ, sizeof(int) * nb));
  HIPCHK(hipMalloc(&d_sf, sizeof(double) * 2 * (size_t)nb));
  HIPCHK(hipMalloc(&d_out, sizeof(double) * 8));
  HIPCHK(hipMemcpyAsync(d_obs, h->obs, sizeof(uint64_t) * nb, hipMemcpyHostToDevice, c->stream));
  int t; rsg_tstart(c, "nullfit", &t);
  hipLaunchKernelGGL(k_nullfit, dim3(1), dim3(1024), 0, c->stream, d_obs, nb, h->imin, h->imax, (double)h->n,
                     (double)h->Nc, h->bmin, h->w, pmass, fracfit, doexpfit, d_S, d_NZ, d_sf, d_out);
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  double o[8];
  double *sf = (double *)malloc(sizeof(double) * 2 * (size_t)nb);
  HIPCHK(hipMemcpyAsync(o, d_out, sizeof(double) * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(sf, d_sf, sizeof(double) * 2 * (size_t)nb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  hipFree(d_obs); hipFree(d_S); hipFree(d_NZ); hipFree(d_sf); hipFree(d_out);
  if (std::isnan(o[4])) { free(sf); return rsg_fail(errbuf, "bad Null histogram fit, pmass is nan."); }
  h->phi = o[5]; h->cmin = (int32_t)o[6]; h->Nc = h->n; h->No = (uint64_t)o[7]; h->tailfit = 1;
  for (int k = 0; k < 5; k++) fit[k] = o[k];
  if (!std::isinf(o[1])) *ret_survfit = sf; else free(sf);
  return RSG_OK;
}

// ---------------------------------------------------------------- E-values
struct EvArgs {
  const unsigned long long *S;   // suffix sums of the null histogram
  const double *sf;              // survfit (2 nb) or null
  double bmin, w, phi, xmax, NcNull, Nc;
  int nb, imin, imax;
};

__device__ double d_cov2evalue(double cov, const EvArgs &A, int *bad)
{
  int icov = (int)ceil(((cov - A.bmin) / A.w) - 1.);
  if (A.sf && icov >= 2 * A.nb - 1) return A.sf[2 * A.nb - 1];
  if (A.sf && cov >= A.phi)         return A.sf[icov + 1];
  if (cov >= A.xmax) return 1.0 / A.NcNull;
  if (icov <= A.imax) {
    if (icov < A.imin) icov = A.imin;
    if (icov >= A.imax - 1) return 1.0 / A.NcNull;
    return (double)A.S[icov] / A.NcNull;
  }
  *bad = 1;
  return NAN;
}

__global__ void k_evalues(const double *__restrict__ cov, int64_t L, EvArgs A, double *__restrict__ ev, int *flag)
{
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= L * L) return;
  int64_t i = idx / L, j = idx - i * L;
  if (i == j) { ev[idx] = INFINITY; return; }
  int bad = 0;
  double p = d_cov2evalue(cov[idx], A, &bad);
  if (bad) atomicOr(flag, 1);
  ev[idx] = p * A.Nc;
}

static int ensure_eval(rsg_ctx *c, rsg_msa *m, char *errbuf)
{
  if (!m->d_eval) HIPCHK(hipMalloc(&m->d_eval, sizeof(double) * (size_t)(m->L * m->L)));
  (void)c;
  return RSG_OK;
}

extern "C" int rsg_cov_Evalues(rsg_ctx *c, rsg_msa *m, const rsg_hist *null, const double *survfit, int64_t Nc,
                               double *Eval, char *errbuf)
{
  if (!m->have_cov) return rsg_fail(errbuf, "no covariation matrix");
  HIPCHK(hipSetDevice(c->device));
  if (ensure_eval(c, m, errbuf) != RSG_OK) return RSG_FAIL;
  const int nb = null->nb;
  const int64_t L = m->L;
  unsigned long long *d_S = nullptr;
  double *d_sf = nullptr;
  int st = RSG_OK;
  // tail counts S[b] = sum_{b' >= b} obs[b'] over [imin, imax] (host: the histogram is here)
  std::vector<unsigned long long> S((size_t)std::max(nb, 1), 0ull);
  unsigned long long run = 0;
  for (int b = null->imax; b >= null->imin && b >= 0; b--) { if (b < nb) { run += null->obs[b]; S[(size_t)b] = run; } }
  if (hipMalloc(&d_S, sizeof(unsigned long long) * S.size()) != hipSuccess ||
      (survfit && hipMalloc(&d_sf, sizeof(double) * 2 * (size_t)nb) != hipSuccess))
    st = rsg_fail(errbuf, "rsg_cov_Evalues: out of device memory");
  if (st == RSG_OK && (hipMemcpyAsync(d_S, S.data(), sizeof(unsigned long long) * S.size(), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
                       (survfit && hipMemcpyAsync(d_sf, survfit, sizeof(double) * 2 * (size_t)nb, hipMemcpyHostToDevice, c->stream) != hipSuccess) ||
                       hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream) != hipSuccess))
    st = rsg_fail(errbuf, "rsg_cov_Evalues: upload failed");
  if (st == RSG_OK) {
    EvArgs A;
    A.S = d_S; A.sf = d_sf; A.bmin = null->bmin; A.w = null->w; A.phi = null->phi; A.xmax = null->xmax;
    A.NcNull = (double)null->Nc; A.Nc = (double)Nc; A.nb = nb; A.imin = null->imin; A.imax = null->imax;
    int t; rsg_tstart(c, "evalues", &t);
    hipLaunchKernelGGL(k_evalues, dim3((unsigned)((L * L + 255) / 256)), dim3(256), 0, c->stream, m->d_cov, L, A, m->d_eval, c->d_flag);
    rsg_tstop(c, t);
    if (hipGetLastError() != hipSuccess) st = rsg_fail(errbuf, "rsg_cov_Evalues: launch failed");
  }
  if (st == RSG_OK && Eval &&
      hipMemcpyAsync(Eval, m->d_eval, sizeof(double) * (size_t)(L * L), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    st = rsg_fail(errbuf, "rsg_cov_Evalues: download failed");
  if (st == RSG_OK && (hipMemcpyAsync(c->h_pinned, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                       hipStreamSynchronize(c->stream) != hipSuccess))
    st = rsg_fail(errbuf, "rsg_cov_Evalues failed");
  if (d_S) hipFree(d_S);
  if (d_sf) hipFree(d_sf);
  if (st == RSG_OK && *(int *)c->h_pinned) st = rsg_fail(errbuf, "cannot find evalue for a covariation score");
  m->have_eval = (st == RSG_OK);
  return st;
}

__global__ void k_evtail(const double *__restrict__ cov, int64_t L, double m, double lambda, double p0, double xcap,
                         double Nc, double *__restrict__ ev)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L * L) return;
  const int64_t i = e / L, j = e - i * L;
  if (i == j) { ev[e] = INFINITY; return; }
  const double p = p0 * exp(-lambda * (fmin(cov[e], xcap) - m));
  ev[e] = Nc * (p < 1.0 ? p : 1.0);
}

extern "C" int rsg_cov_EvaluesTail(rsg_ctx *c, rsg_msa *m, double mu, double lambda, double p0, double xcap,
                                   int64_t Nc, double *Eval, char *errbuf)
{
  if (!c || !m || std::isnan(xcap)) return rsg_fail(errbuf, "rsg_cov_EvaluesTail: bad arguments");
  if (!m->have_cov) return rsg_fail(errbuf, "no covariation matrix");
  if (!(lambda > 0.) || !(p0 > 0.)) return rsg_fail(errbuf, "rsg_cov_EvaluesTail: needs lambda > 0 and p0 > 0");
  HIPCHK(hipSetDevice(c->device));
  if (ensure_eval(c, m, errbuf) != RSG_OK) return RSG_FAIL;
  const int64_t L = m->L;
  int t; rsg_tstart(c, "evalues_tail", &t);
  hipLaunchKernelGGL(k_evtail, dim3((unsigned)((L * L + 255) / 256)), dim3(256), 0, c->stream, m->d_cov, L, mu, lambda,
                     p0, xcap, (double)Nc, m->d_eval);
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  if (Eval) HIPCHK(hipMemcpyAsync(Eval, m->d_eval, sizeof(double) * (size_t)(L * L), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  m->have_eval = 1;
  return RSG_OK;
}

// Pairs i < j with E <= thresh, compacted in the order of an atomic counter (the host sorts).
__global__ void k_hits(const double *__restrict__ ev, const double *__restrict__ cov, int64_t L, double thresh,
                       int cap, int *__restrict__ count, int2 *__restrict__ ij, double2 *__restrict__ se)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L * L) return;
  const int64_t i = e / L, j = e - i * L;
  if (i >= j || !(ev[e] <= thresh)) return;
  const int k = atomicAdd(count, 1);
  if (k < cap) { ij[k] = make_int2((int)i, (int)j); se[k] = make_double2(cov[e], ev[e]); }
}

extern "C" int rsg_cov_Hits(rsg_ctx *c, rsg_msa *m, double thresh, int cap, int *nhits, int *hi, int *hj,
                            double *score, double *evalue, char *errbuf)
{
  if (!c || !m || !nhits || cap < 0 || (cap > 0 && (!hi || !hj || !score || !evalue)))
    return rsg_fail(errbuf, "rsg_cov_Hits: bad arguments");
  if (!m->have_eval) return rsg_fail(errbuf, "rsg_cov_Hits: no E-values (run rsg_cov_Evalues or rsg_cov_EvaluesTail)");
  HIPCHK(hipSetDevice(c->device));
  const int64_t L = m->L;
  const size_t dcap = (size_t)std::max(cap, 1);
  if (c->hits_cap < dcap) {                      // grow-only: repeated calls allocate nothing
    if (c->d_hits) hipFree(c->d_hits);
    if (c->h_hits) hipHostFree(c->h_hits);
    c->d_hits = c->h_hits = nullptr; c->hits_cap = 0;
    HIPCHK(hipMalloc(&c->d_hits, (sizeof(double2) + sizeof(int2)) * dcap));
    HIPCHK(hipHostMalloc(&c->h_hits, (sizeof(double2) + sizeof(int2)) * dcap, hipHostMallocDefault));
    c->hits_cap = dcap;
  }
  double2 *d_se = (double2 *)c->d_hits, *se = (double2 *)c->h_hits;
  int2 *d_ij = (int2 *)(d_se + c->hits_cap), *ij = (int2 *)(se + c->hits_cap);
  int st = RSG_OK;
  int *d_count = (int *)c->d_flag + 8;
  // the count and the first `pre` entries come back with one synchronisation; more hits than
  // that take a second copy
  const int pre = std::min(cap, 4096);
  if (hipMemsetAsync(d_count, 0, sizeof(int), c->stream) != hipSuccess) st = rsg_fail(errbuf, "rsg_cov_Hits: memset failed");
  if (st == RSG_OK) {
    hipLaunchKernelGGL(k_hits, dim3((unsigned)((L * L + 255) / 256)), dim3(256), 0, c->stream, m->d_eval, m->d_cov, L, thresh,
                       cap, d_count, d_ij, d_se);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(c->h_pinned, d_count, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        (pre > 0 && (hipMemcpyAsync(ij, d_ij, sizeof(int2) * pre, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                     hipMemcpyAsync(se, d_se, sizeof(double2) * pre, hipMemcpyDeviceToHost, c->stream) != hipSuccess)) ||
        hipStreamSynchronize(c->stream) != hipSuccess) st = rsg_fail(errbuf, "rsg_cov_Hits failed");
  }
  int n = (st == RSG_OK) ? *(int *)c->h_pinned : 0;
  const int got = std::min(n, cap);
  if (st == RSG_OK && got > pre) {
    if (hipMemcpyAsync(ij + pre, d_ij + pre, sizeof(int2) * (got - pre), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(se + pre, d_se + pre, sizeof(double2) * (got - pre), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) st = rsg_fail(errbuf, "rsg_cov_Hits: download failed");
  }
  if (st != RSG_OK) return st;
  // the atomic order is arbitrary: sort by (i, j)
  std::vector<int> ord((size_t)got);
  for (int k = 0; k < got; k++) ord[(size_t)k] = k;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) {
    return ij[(size_t)a].x != ij[(size_t)b].x ? ij[(size_t)a].x < ij[(size_t)b].x : ij[(size_t)a].y < ij[(size_t)b].y;
  });
  for (int k = 0; k < got; k++) {
    const int q = ord[(size_t)k];
    hi[k] = ij[(size_t)q].x; hj[k] = ij[(size_t)q].y; score[k] = se[(size_t)q].x; evalue[k] = se[(size_t)q].y;
  }
  *nhits = n;
  return (n > cap) ? rsg_fail(errbuf, "rsg_cov_Hits: %d hits exceed the capacity %d (the first %d returned)", n, cap, cap) : RSG_OK;
}

// ---------------------------------------------------------------- scalar helpers (host)
extern "C" double rsg_cov2evalue(double cov, int64_t Nc, const rsg_hist *h, const double *survfit)
{
  int icov;
  if (score2bin(h, cov, &icov) != RSG_OK) return NAN;
  if (survfit && icov >= 2 * h->nb - 1) return survfit[2 * h->nb - 1] * (double)Nc;
  if (survfit && cov >= h->phi)         return survfit[icov + 1] * (double)Nc;
  if (cov >= h->xmax) return (double)Nc / (double)h->Nc;
  if (icov <= h->imax) {
    if (icov < h->imin) icov = h->imin;
    if (icov >= h->imax - 1) return (double)Nc / (double)h->Nc;
    uint64_t cc = 0;
    for (int i = h->imax; i >= icov; i--) cc += h->obs[i];
    return (double)cc * (double)Nc / (double)h->Nc;
  }
  return NAN;
}

extern "C" double rsg_evalue2cov(double eval_thresh, int64_t Nc, const rsg_hist *h, const double *survfit)
{
  double cov = -INFINITY, eval;
  int b = h->cmin - 1, i;
  if (h->No >= 10 && survfit) {
    for (b = 2 * h->nb - 1; b >= h->cmin; b--) {
      eval = survfit[b] * (double)Nc;
      if (eval >= eval_thresh) break;
    }
    cov = lbound(h, b);
  }
  if (b == h->cmin - 1) {
    uint64_t cc = 0;
    for (i = h->imax; i >= h->imin; i--) {
      cc += h->obs[i];
      eval = (double)cc * (double)Nc / (double)h->Nc;
      if (eval >= eval_thresh) break;
    }
    cov = ubound(h, i);
  }
  return cov;
}

// ---------------------------------------------------------------- .null text format
// cov_WriteNullHistogram: "%f %llu\n" per bin, x = bmin + i w (covariation.c:1845-1868).
extern "C" int rsg_hist_WriteNull(const char *path, const rsg_hist *h, char *errbuf)
{
  FILE *fp = fopen(path, "w");
  if (!fp) return rsg_fail(errbuf, "could not open nullhisfile %s", path);
  for (int i = 0; i < h->nb; i++) fprintf(fp, "%f %llu\n", h->bmin + i * h->w, (unsigned long long)h->obs[i]);
  fclose(fp);
  return RSG_OK;
}

// cov_ReadNullHistogram (covariation.c:1717-1843): w = smallest gap between successive
// scores, bmin / bmax = score extremes, then every "score count" line sets its bin.
extern "C" int rsg_hist_ReadNull(const char *path, rsg_hist **ret, char *errbuf)
{
  *ret = nullptr;
  FILE *fp = fopen(path, "r");
  if (!fp) return rsg_fail(errbuf, "file open failed");
  std::vector<double> sc; std::vector<unsigned long long> cnt;
  char line[1024];
  while (fgets(line, sizeof line, fp)) {
    char *p = line;
    while (*p == ' ' || *p == '\t') p++;
    if (*p == '#' || *p == '\n' || *p == 0) continue;
    // every line's first token is a score (it enters w and the range); only a line of exactly
    // two tokens sets a bin, to its last token (atoi)
    int t = 0;
    unsigned long long c = 0ull;
    for (char *tok = strtok(p, " \t\n"); tok; tok = strtok(nullptr, " \t\n"), t++) {
      if (t == 0) sc.push_back(atof(tok));
      else        c = (unsigned long long)atoi(tok);
    }
    if (t == 0) continue;
    cnt.push_back(t == 2 ? c : ~0ull);                      // ~0: a line that sets no bin
  }
  fclose(fp);
  if (sc.empty()) return rsg_fail(errbuf, "empty null histogram file");
  double w = INFINITY, prv = -INFINITY, mx = -INFINITY, mn = INFINITY;
  for (double s : sc) { w = std::min(w, fabs(s - prv)); prv = s; mx = std::max(mx, s); mn = std::min(mn, s); }
  double bmax = mx, bmin = mn;
  while (fabs(bmax - bmin) < 1e-2) bmax += w;
  rsg_hist *h = rsg_hist_create(bmin, bmax, w);
  for (size_t k = 0; k < sc.size(); k++) {
    if (cnt[k] == ~0ull) continue;
    int b;
    if (score2bin(h, sc[k], &b) != RSG_OK) { rsg_hist_destroy(h); return rsg_fail(errbuf, "bad score %f", sc[k]); }
    b = grow_to(h, b);
    if (b < h->imin) { h->imin = b; h->cmin = b; }
    if (b > h->imax) h->imax = b;
    if (sc[k] < h->xmin) h->xmin = sc[k];
    if (sc[k] > h->xmax) h->xmax = sc[k];
    h->obs[b] = cnt[k];
    h->n += cnt[k]; h->Nc += cnt[k]; h->No += cnt[k];
  }
  *ret = h;
  return RSG_OK;
}
