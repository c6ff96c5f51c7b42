// rsg_capi.hip -- extern "C" entry points of librscape_hip (declared in include/rscape_hip.h).
#include "rsg_internal.h"
#include <cmath>
#include <chrono>
#include <cstring>
#include <new>

static const double kAllowWCGU[16] = {   // R-scape default allowpair: WC + GU (src/R-scape.c:882-887)
  0, 0, 0, 1,
  0, 0, 1, 0,
  0, 1, 0, 1,
  1, 0, 1, 0 };

extern "C" rsg_ctx *rsg_ctx_create(int device, char *errbuf)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device) { rsg_fail(errbuf, "no HIP device %d (count %d)", device, n); return nullptr; }
  if (hipSetDevice(device) != hipSuccess) { rsg_fail(errbuf, "hipSetDevice(%d) failed", device); return nullptr; }
  rsg_ctx *c = new (std::nothrow) rsg_ctx();
  if (!c) { rsg_fail(errbuf, "out of memory"); return nullptr; }
  c->device = device;
  hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->d_flag, 64) != hipSuccess || hipMalloc(&c->d_small, 65536) != hipSuccess ||
      hipMalloc(&c->d_res, 8 * sizeof(double)) != hipSuccess ||
      hipMemset(c->d_res, 0, 8 * sizeof(double)) != hipSuccess ||
      hipHostMalloc((void **)&c->h_pinned, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
      hipHostGetDevicePointer((void **)&c->d_hres, c->h_pinned, 0) != hipSuccess) {
    rsg_fail(errbuf, "context allocation failed");
    delete c;
    return nullptr;
  }
  memset(c->h_pinned, 0, 64 * sizeof(double));          // completion stamps start from 1
  if (rsg_noalloc_init(device, errbuf) != RSG_OK) { rsg_ctx_destroy(c); return nullptr; }
  return c;
}

extern "C" void rsg_ctx_destroy(rsg_ctx *c)
{
  if (!c) return;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  for (int k = 0; k < RSG_MAX_TIMERS; k++)
    if (c->timers[k].a) { hipEventDestroy(c->timers[k].a); hipEventDestroy(c->timers[k].b); }
  if (c->d_red) hipFree(c->d_red);
  if (c->d_gscr) hipFree(c->d_gscr);
  if (c->bexec) hipGraphExecDestroy(c->bexec);
  if (c->d_arena) hipFree(c->d_arena);
  if (c->d_hits) hipFree(c->d_hits);
  if (c->h_hits) hipHostFree(c->h_hits);
  if (c->d_flag) hipFree(c->d_flag);
  if (c->d_res) hipFree(c->d_res);
  if (c->d_small) hipFree(c->d_small);
  if (c->h_pinned) hipHostFree(c->h_pinned);
  hipStreamDestroy(c->stream);
  delete c;
}

__global__ void __launch_bounds__(256) k_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n)
{
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) dst[k] = src[k];
}

extern "C" int rsg_ctx_copy_probe(rsg_ctx *c, size_t bytes, char *errbuf)
{
  if (!c || bytes == 0 || bytes % 16) return rsg_fail(errbuf, "rsg_ctx_copy_probe: bytes must be a positive multiple of 16");
  HIPCHK(hipSetDevice(c->device));
  uint4 *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, bytes));
  if (hipMalloc(&b, bytes) != hipSuccess) { hipFree(a); return rsg_fail(errbuf, "rsg_ctx_copy_probe: out of device memory"); }
  int st = RSG_OK;
  if (hipMemsetAsync(a, 1, bytes, c->stream) != hipSuccess) st = rsg_fail(errbuf, "rsg_ctx_copy_probe: memset failed");
  const int64_t n = (int64_t)(bytes / 16);
  for (int r = 0; r < 3 && st == RSG_OK; r++) {
    int t; rsg_tstart(c, "copy_probe", &t);
    hipLaunchKernelGGL(k_copy_probe, dim3(2048), dim3(256), 0, c->stream, a, b, n);
    rsg_tstop(c, t);
    if (hipGetLastError() != hipSuccess) st = rsg_fail(errbuf, "rsg_ctx_copy_probe: launch failed");
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess && st == RSG_OK) st = rsg_fail(errbuf, "rsg_ctx_copy_probe: failed");
  hipFree(a); hipFree(b);
  return st;
}

extern "C" int rsg_ctx_set_timing(rsg_ctx *c, int on) { c->timing = on; c->ntimers = 0; return RSG_OK; }

extern "C" int rsg_ctx_timings(rsg_ctx *c, const char **names, double *ms, int *n, int cap)
{
  hipStreamSynchronize(c->stream);
  int k = 0;
  for (; k < c->ntimers && k < cap; k++) {
    float f = 0.f;
    if (c->timers[k].host_ms >= 0.) f = (float)c->timers[k].host_ms;
    else hipEventElapsedTime(&f, c->timers[k].a, c->timers[k].b);
    names[k] = c->timers[k].name; ms[k] = f;
  }
  *n = k;
  c->ntimers = 0;
  return RSG_OK;
}

extern "C" int rsg_ctx_set_gram_shape(rsg_ctx *c, int shape) { c->gram_shape = shape; return RSG_OK; }
extern "C" int rsg_ctx_set_gram_split(rsg_ctx *c, int on) { c->gram_split = on ? 1 : 0; return RSG_OK; }
extern "C" int rsg_ctx_set_batch_persistent(rsg_ctx *c, int on) { c->batch_persistent = on ? 1 : 0; return RSG_OK; }

static void free_msa(rsg_msa *m)
{
  void *ptrs[] = { m->d_ax, m->d_axT, m->d_xt, m->d_w, m->d_wsplit, m->d_wq, m->d_nsx,
                   m->d_cnt, m->d_cnt_u, m->d_pm, m->d_ps, m->d_lgc, m->d_cov, m->d_tmp, m->d_rowx, m->d_eval };
  for (void *p : ptrs) if (p) hipFree(p);
  delete m;
}

extern "C" int rsg_msa_update(rsg_ctx *c, rsg_msa *m, const uint8_t *ax, const double *wgt, char *errbuf)
{
  const size_t nl = (size_t)m->N * (size_t)m->L;
  if (ax) {
    HIPCHK(hipMemcpyAsync(m->d_ax, ax, nl, hipMemcpyHostToDevice, c->stream));
    if (rsg_pack(c, m, errbuf) != RSG_OK) return RSG_FAIL;
  }
  std::vector<double> w((size_t)m->N, 1.0);
  if (wgt) memcpy(w.data(), wgt, sizeof(double) * (size_t)m->N);
  m->weighted = (wgt != nullptr);
  m->wsum = 0.;
  m->weights_ok = 1;
  for (double v : w) { m->wsum += v; if (!(v >= 0.) || !std::isfinite(v)) m->weights_ok = 0; }
  {                                 // the rounding of k_quant_weights, repeated for the rounded total
    double wmax = 0.;
    for (double v : w) wmax = std::fmax(wmax, v);
    const double qmax = (double)((1 << 21) - 1), W = (wmax > 0.) ? wmax / qmax : 1.;
    m->wsum_q = 0.;
    for (double v : w) {
      double q = (wmax > 0.) ? std::rint(v / W) : 0.;
      q = std::fmin(std::fmax(q, 0.), qmax);
      m->wsum_q += W * q;
    }
  }
  HIPCHK(hipMemcpyAsync(m->d_w, w.data(), sizeof(double) * (size_t)m->N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  m->nsplit_ready = 0;
  m->wq_ready = 0;
  m->have_probs = m->have_cov = m->have_eval = 0;
  return RSG_OK;
}

// a resident alignment's buffers (codes, packed panels, weights), contents unset
int rsg_msa_alloc(rsg_ctx *c, int64_t nseq, int64_t alen, rsg_msa **ret, char *errbuf)
{
  *ret = nullptr;
  if (nseq < 1 || alen < 1) return rsg_fail(errbuf, "empty alignment (nseq %lld, alen %lld)", (long long)nseq, (long long)alen);
  if (nseq > (1ll << 24)) return rsg_fail(errbuf, "nseq %lld exceeds the int32 one-hot Gram range", (long long)nseq);
  HIPCHK(hipSetDevice(c->device));
  rsg_msa *m = new (std::nothrow) rsg_msa();
  if (!m) return rsg_fail(errbuf, "out of memory");
  m->N = nseq; m->L = alen;
  m->Npad = rsg_npad(nseq);
  m->Lpad = ((alen + 63) / 64) * 64;
  const size_t nl = (size_t)nseq * (size_t)alen, tp = (size_t)m->Lpad * (size_t)m->Npad;
  if (hipMalloc(&m->d_ax, nl) != hipSuccess || hipMalloc(&m->d_axT, tp) != hipSuccess ||
      hipMalloc(&m->d_xt, tp) != hipSuccess || hipMalloc(&m->d_w, sizeof(double) * (size_t)nseq) != hipSuccess) {
    free_msa(m);
    return rsg_fail(errbuf, "device allocation failed for a %lld x %lld alignment", (long long)nseq, (long long)alen);
  }
  *ret = m;
  return RSG_OK;
}

extern "C" int rsg_msa_create(rsg_ctx *c, const uint8_t *ax, const double *wgt, int64_t nseq, int64_t alen,
                              rsg_msa **ret, char *errbuf)
{
  rsg_msa *m = nullptr;
  if (rsg_msa_alloc(c, nseq, alen, &m, errbuf) != RSG_OK) return RSG_FAIL;
  if (rsg_msa_update(c, m, ax, wgt, errbuf) != RSG_OK) { free_msa(m); return RSG_FAIL; }
  *ret = m;
  return RSG_OK;
}

extern "C" void rsg_msa_destroy(rsg_ctx *c, rsg_msa *m)
{
  if (!m) return;
  if (c) hipStreamSynchronize(c->stream);
  free_msa(m);
}

static int alloc_counts(rsg_msa *m, void **dst, int kind, char *errbuf)
{
  const size_t bytes = (size_t)rsg_rec_bytes(kind) * (size_t)m->L * (size_t)m->L;
  if (*dst && m->cnt_bytes >= bytes && dst == &m->d_cnt) return RSG_OK;
  if (*dst) hipFree(*dst);
  *dst = nullptr;
  HIPCHK(hipMalloc(dst, bytes));
  if (dst == &m->d_cnt) m->cnt_bytes = bytes;
  return RSG_OK;
}

extern "C" int rsg_corr_Probs(rsg_ctx *c, rsg_msa *m, int mode, double tol, char *errbuf)
{
  HIPCHK(hipSetDevice(c->device));
  m->have_probs = m->have_cov = m->have_eval = 0;
  m->wq_counts = 0;
  if (mode == RSG_COUNTS_INT8) {
    if (m->weighted) return rsg_fail(errbuf, "int8 counts need unit weights (alignment has weights)");
    const int kind = rsg_int_kind(m->N);
    if (alloc_counts(m, &m->d_cnt, kind, errbuf) != RSG_OK) return RSG_FAIL;
    m->cnt_kind = kind;
    double *pm_part = nullptr;                    // uint16 records: marginal partials from the Gram epilogue
    if (kind == CNT_U16 && m->L >= 2 && !(pm_part = rsg_pm32_scratch(c, m->L, errbuf))) return RSG_FAIL;
    m->pm_fused = 0;
    if (rsg_gram_counts_i8(c, m, m->d_cnt, kind, pm_part, (int *)(c->d_res + 4), errbuf) != RSG_OK) return RSG_FAIL;
    m->pm_fused = (pm_part != nullptr);
  } else if (mode == RSG_COUNTS_W8) {
    if (!m->weights_ok) return rsg_fail(errbuf, "pp validation failed: sequence weights must be finite and >= 0");
    m->pm_fused = 0;
    if (alloc_counts(m, &m->d_cnt, CNT_F32, errbuf) != RSG_OK) return RSG_FAIL;
    m->cnt_kind = CNT_F32;
    if (rsg_gram_counts_w8(c, m, (float *)m->d_cnt, errbuf) != RSG_OK) return RSG_FAIL;
    m->wq_counts = 1;
  } else if (mode == RSG_COUNTS_BF16 || mode == RSG_COUNTS_BF16X3) {
    // a negative or non-finite weight makes pp fail corr_ValidateProbs (src/correlators.c:1499-1545)
    if (!m->weights_ok) return rsg_fail(errbuf, "pp validation failed: sequence weights must be finite and >= 0");
    m->pm_fused = 0;
    if (alloc_counts(m, &m->d_cnt, CNT_F32, errbuf) != RSG_OK) return RSG_FAIL;
    m->cnt_kind = CNT_F32;
    if (rsg_gram_counts_bf16(c, m, mode == RSG_COUNTS_BF16X3 ? 3 : 2, (float *)m->d_cnt, errbuf) != RSG_OK) return RSG_FAIL;
  } else return rsg_fail(errbuf, "unknown counts mode %d", mode);
  if (rsg_probs_from_counts(c, m, tol, errbuf) != RSG_OK) return RSG_FAIL;
  m->have_probs = 1;
  if (c->defer) return RSG_OK;                 // checked by the deferring caller
  HIPCHK(hipMemcpyAsync(c->h_pinned + 4, c->d_res + 4, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (*(int *)(c->h_pinned + 4)) { m->have_probs = 0; return rsg_fail(errbuf, "pm validation failed"); }
  return RSG_OK;
}

extern "C" int rsg_corr_GetCounts(rsg_ctx *c, rsg_msa *m, double *cnt, char *errbuf)
{
  if (!m->d_cnt) return rsg_fail(errbuf, "no counts");
  const int64_t L = m->L;
  const size_t bytes = (size_t)rsg_rec_bytes(m->cnt_kind) * (size_t)L * (size_t)L;
  std::vector<uint8_t> h(bytes);
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(h.data(), m->d_cnt, bytes, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < L; i++)
    for (int64_t j = 0; j < L; j++)
      for (int k = 0; k < 16; k++) {
        size_t o = ((size_t)i * L + j) * 16 + k;
        size_t r = (size_t)rsg_rec(i, j, L) * 16 + k;
        double v = 0.;
        if (i < j) {
          if (m->cnt_kind == CNT_U16) v = (double)((const uint16_t *)h.data())[r];
          else if (m->cnt_kind == CNT_U32) v = (double)((const uint32_t *)h.data())[r];
          else                        v = ((float *)h.data())[r];
        }
        cnt[o] = v;
      }
  return RSG_OK;
}

extern "C" int rsg_corr_GetProbs(rsg_ctx *c, rsg_msa *m, double *pp, double *pm, double *ps,
                                 double *nseff, double *ngap, char *errbuf)
{
  if (!m->have_probs) return rsg_fail(errbuf, "corr_Probs has not run");
  const int64_t L = m->L;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (pm) HIPCHK(hipMemcpy(pm, m->d_pm, sizeof(double) * 4 * (size_t)L, hipMemcpyDeviceToHost));
  if (ps) HIPCHK(hipMemcpy(ps, m->d_ps, sizeof(double) * 5 * (size_t)L, hipMemcpyDeviceToHost));
  if (pp || nseff || ngap) {
    std::vector<double> cnt(16 * (size_t)L * (size_t)L);
    if (rsg_corr_GetCounts(c, m, cnt.data(), errbuf) != RSG_OK) return RSG_FAIL;
    for (int64_t i = 0; i < L; i++)
      for (int64_t j = i + 1; j < L; j++) {
        const double *r = &cnt[((size_t)i * L + j) * 16];
        double ns = 0., tot = 0.;
        for (int k = 0; k < 16; k++) { ns += r[k]; tot += r[k] + 1e-10; }
        if (nseff) nseff[i * L + j] = nseff[j * L + i] = ns;
        if (ngap) ngap[i * L + j] = rsg_count_wsum(m) - ns;
        if (pp)
          for (int x = 0; x < 4; x++)
            for (int y = 0; y < 4; y++) {
              double v = (r[x * 4 + y] + 1e-10) / tot;
              pp[((size_t)i * L + j) * 16 + x * 4 + y] = v;
              pp[((size_t)j * L + i) * 16 + y * 4 + x] = v;
            }
      }
    for (int64_t i = 0; i < L; i++) {
      if (nseff) nseff[i * L + i] = 0.;
      if (ngap) ngap[i * L + i] = 0.;
      if (pp) for (int k = 0; k < 16; k++) pp[((size_t)i * L + i) * 16 + k] = 0.;
    }
  }
  return RSG_OK;
}

static int resolve_class(const rsg_msa *m, int cls, int nseqthresh, int alenthresh)
{
  if (cls == RSG_CSELECT) return (m->N <= nseqthresh || m->L <= alenthresh) ? RSG_C2 : RSG_C16;
  return cls;
}

extern "C" int rsg_corr_CalculateCOV(rsg_ctx *c, rsg_msa *m, int covtype, int covclass, const double *allowpair,
                                     int nseqthresh, int alenthresh, int prec, char *errbuf)
{
  HIPCHK(hipSetDevice(c->device));
  const double *allow = allowpair ? allowpair : kAllowWCGU;
  int base = covtype - covtype % 3;
  if (base == RSG_RAF || base == RSG_RAFS)
    return rsg_corr_CalculateRAF(c, m, base == RSG_RAFS, allow, prec, errbuf);
  if (!m->have_probs) return rsg_fail(errbuf, "corr_Probs has not run");
  int cls = resolve_class(m, covclass, nseqthresh, alenthresh);
  if (base == RSG_CCF) cls = RSG_C16;
  if (cls == RSG_CWC && base != RSG_GT) return rsg_fail(errbuf, "CWC not implemented for covariation type %d", base);
  if (cls < RSG_C16 || cls > RSG_CWC) return rsg_fail(errbuf, "bad covariation class %d", covclass);
  return rsg_stat(c, m, base, cls, allow, prec, errbuf);
}

extern "C" int rsg_corr_CalculateRAF(rsg_ctx *c, rsg_msa *m, int rafs, const double *allowpair, int prec, char *errbuf)
{
  (void)prec;
  HIPCHK(hipSetDevice(c->device));
  const double *allow = allowpair ? allowpair : kAllowWCGU;
  // RAF is defined on unweighted symbols (src/correlators.c:892-921) and, like the reference,
  // does not go through corr_Probs: form the unit-weight counts here.
  m->cnt_u_kind = rsg_int_kind(m->N);
  if (!m->d_cnt_u) HIPCHK(hipMalloc(&m->d_cnt_u, (size_t)rsg_rec_bytes(m->cnt_u_kind) * (size_t)m->L * (size_t)m->L));
  if (rsg_gram_counts_i8(c, m, m->d_cnt_u, m->cnt_u_kind, nullptr, nullptr, errbuf) != RSG_OK) return RSG_FAIL;
  return rsg_raf(c, m, rafs, allow, errbuf);
}

extern "C" int rsg_corr_CalculateCOVCorrected(rsg_ctx *c, rsg_msa *m, int actype, int shiftnonneg, char *errbuf)
{
  HIPCHK(hipSetDevice(c->device));
  if (actype != RSG_APC && actype != RSG_ASC) return rsg_fail(errbuf, "wrong correction type");
  return rsg_correct(c, m, actype, shiftnonneg, errbuf);
}

extern "C" int rsg_corr_GetCOV(rsg_ctx *c, rsg_msa *m, double *COV, double *minCOV, double *maxCOV, char *errbuf)
{
  if (!m->have_cov) return rsg_fail(errbuf, "no covariation matrix");
  HIPCHK(hipStreamSynchronize(c->stream));
  if (COV) HIPCHK(hipMemcpy(COV, m->d_cov, sizeof(double) * (size_t)m->L * (size_t)m->L, hipMemcpyDeviceToHost));
  if (minCOV) *minCOV = m->minCOV;
  if (maxCOV) *maxCOV = m->maxCOV;
  return RSG_OK;
}

// Wait for the completion stamp of the chain's last kernel in pinned memory (the result block is
// complete once it shows); true when it was seen.  The stream is polled between spins, so a chain
// that ends without writing the stamp (a failed launch) falls through to the caller's stream
// synchronisation, which reports the failure; so does a chain still running after
// stamp_wait_us (2 ms by default: large alignments), which then waits there instead of keeping
// a core spinning.  stamp_wait_us = 0 skips the spin.
static bool wait_stamp(rsg_ctx *c)
{
  if (c->stamp_wait_us <= 0) return false;
  const unsigned long long want = c->hstamp;
  unsigned long long *slot = (unsigned long long *)(c->h_pinned + 63);
  const auto t0 = std::chrono::steady_clock::now();
  const auto limit = std::chrono::microseconds(c->stamp_wait_us);
  for (unsigned spin = 1;; spin++) {
    if (__atomic_load_n(slot, __ATOMIC_ACQUIRE) == want) return true;
    if ((spin & 255) == 0) {
      if (hipStreamQuery(c->stream) != hipErrorNotReady) return __atomic_load_n(slot, __ATOMIC_ACQUIRE) == want;
      if (std::chrono::steady_clock::now() - t0 > limit) return false;
    }
  }
}

extern "C" int rsg_ctx_set_stamp_wait(rsg_ctx *c, int microseconds)
{
  c->stamp_wait_us = microseconds < 0 ? 0 : microseconds;
  return RSG_OK;
}

extern "C" int rsg_cov_Calculate(rsg_ctx *c, rsg_msa *m, int covtype, int covclass, int counts_mode, int prec,
                                 const double *allowpair, int nseqthresh, int alenthresh, double tol,
                                 double *minCOV, double *maxCOV, char *errbuf)
{
  int base = covtype - covtype % 3, corr = covtype % 3;
  // the whole chain is queued with one synchronisation at the end: the kernels leave the pm
  // validation flag, the correction's NaN flag, COVavg and the final min / max in d_res, and
  // one copy brings them to h_pinned
  if (rsg_reserve_scratch(c, m->L, errbuf) != RSG_OK) return RSG_FAIL;
  c->defer = 1;
  c->skip_minmax = (corr != 0);
  int st = RSG_OK;
  if (base != RSG_RAF && base != RSG_RAFS) st = rsg_corr_Probs(c, m, counts_mode, tol, errbuf);
  if (st == RSG_OK) st = rsg_corr_CalculateCOV(c, m, base, covclass, allowpair, nseqthresh, alenthresh, prec, errbuf);
  c->skip_minmax = 0;
  if (st == RSG_OK && corr == 1) st = rsg_corr_CalculateCOVCorrected(c, m, RSG_APC, 0, errbuf);
  if (st == RSG_OK && corr == 2) st = rsg_corr_CalculateCOVCorrected(c, m, RSG_ASC, 0, errbuf);
  c->defer = 0;
  if (st == RSG_OK && !c->res_on_host &&
      hipMemcpyAsync(c->h_pinned, c->d_res, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    st = rsg_fail(errbuf, "rsg_cov_Calculate: result copy failed");
  const bool stamped = (st == RSG_OK && c->res_on_host);
  c->res_on_host = 0;
  if (!(stamped && wait_stamp(c)) && hipStreamSynchronize(c->stream) != hipSuccess && st == RSG_OK)
    st = rsg_fail(errbuf, "rsg_cov_Calculate: device failure");
  if (st != RSG_OK) { m->have_probs = m->have_cov = 0; return st; }
  if (base != RSG_RAF && base != RSG_RAFS && *(int *)(c->h_pinned + 4)) {
    m->have_probs = m->have_cov = 0;
    return rsg_fail(errbuf, "pm validation failed");
  }
  if (corr && *(int *)(c->h_pinned + 2)) { m->have_cov = 0; return rsg_fail(errbuf, "bad covariation COVavg %f", c->h_pinned[3]); }
  m->minCOV = c->h_pinned[0]; m->maxCOV = c->h_pinned[1];
  if (minCOV) *minCOV = m->minCOV;
  if (maxCOV) *maxCOV = m->maxCOV;
  return RSG_OK;
}

extern "C" void rsg_free(void *p) { free(p); }
