// rsg_device.h -- device helpers shared by the single-alignment and batched kernels.
#pragma once
#include "rsg_internal.h"
#include <cmath>

// A record as raw words (16-byte loads issued ahead of their use) and its decoding.
template <int KIND> struct RawRec { uint4 q[KIND == CNT_U16 ? 2 : 4]; };
template <int KIND> __device__ __forceinline__ void load_raw(const void *cnt, int64_t rec, RawRec<KIND> &r)
{
  const uint4 *p = (const uint4 *)((const uint8_t *)cnt + rec * (KIND == CNT_U16 ? 32 : 64));
#pragma unroll
  for (int h = 0; h < (KIND == CNT_U16 ? 2 : 4); h++) r.q[h] = p[h];
}
template <int KIND> __device__ __forceinline__ void decode_rec(const RawRec<KIND> &r, double *c)
{
  if (KIND == CNT_U16) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const unsigned w[4] = {r.q[h].x, r.q[h].y, r.q[h].z, r.q[h].w};
#pragma unroll
      for (int q = 0; q < 4; q++) { c[8 * h + 2 * q] = (double)(w[q] & 0xFFFFu); c[8 * h + 2 * q + 1] = (double)(w[q] >> 16); }
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint4 v = r.q[k];
      if (KIND == CNT_U32) { c[4 * k] = v.x; c[4 * k + 1] = v.y; c[4 * k + 2] = v.z; c[4 * k + 3] = v.w; }
      else { c[4 * k] = __uint_as_float(v.x); c[4 * k + 1] = __uint_as_float(v.y); c[4 * k + 2] = __uint_as_float(v.z); c[4 * k + 3] = __uint_as_float(v.w); }
    }
  }
}

template <int KIND> __device__ __forceinline__ void load_rec(const void *cnt, int64_t rec, double *c)
{
  if (KIND == CNT_U16) {
    const uint4 *p = (const uint4 *)((const uint16_t *)cnt + rec * 16);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const uint4 v = p[h];                         // cells 8h .. 8h+7, two per word
      const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int q = 0; q < 4; q++) { c[8 * h + 2 * q] = (double)(w[q] & 0xFFFFu); c[8 * h + 2 * q + 1] = (double)(w[q] >> 16); }
    }
  } else if (KIND == CNT_U32) {
    const uint4 *p = (const uint4 *)((const uint32_t *)cnt + rec * 16);
#pragma unroll
    for (int k = 0; k < 4; k++) { const uint4 v = p[k]; c[4 * k] = v.x; c[4 * k + 1] = v.y; c[4 * k + 2] = v.z; c[4 * k + 3] = v.w; }
  } else {
    const float4 *p = (const float4 *)((const float *)cnt + rec * 16);
#pragma unroll
    for (int k = 0; k < 4; k++) { float4 v = p[k]; c[4 * k] = v.x; c[4 * k + 1] = v.y; c[4 * k + 2] = v.z; c[4 * k + 3] = v.w; }
  }
}


// per-pair statistics
struct StatArgs {
  const void *cnt; const double *pm; const double *lpm; const double *lgc;
  int64_t L, N; double wsum; double *cov; double *part;  // part: 2 per block (min, max)
  int type, cls; unsigned allow;                           // allow: 16-bit mask of allowed cells
  double meanp[4];                                         // CCF
  const double *nsx;                                       // nseff per record (FP64), or nullptr: sum of the cells
};

template <typename T> __device__ __forceinline__ T tlog(T v);
template <> __device__ __forceinline__ double tlog<double>(double v) { return log(v); }
template <> __device__ __forceinline__ float tlog<float>(float v) { return logf(v); }

// phi(r) = r ln r - r + 1 >= 0 for r = p / q > 0, from FP32 values of p, q and d = p - q (d formed
// in FP64 from the unrounded values, so it carries no cancellation).  Near r = 1 the direct form
// cancels, so for |v| < 1/4, v = (p - q) / (p + q), it uses Loader's series (C. Loader, "Fast and
// accurate computation of binomial probabilities", 2000: x ln(x/m) - x + m = (x - m) v +
// 2 x sum_k v^(2k+1) / (2k + 1)) through v^13 (truncation < 2e-8 relative); elsewhere the direct
// form loses at most a factor 4.6 to cancellation.  Reciprocals and the logarithm are the
// hardware's (v_rcp_f32, v_log_f32: about 1 ulp); both branches are evaluated (no divergence).
__device__ __forceinline__ float phi_pq(float pf, float qf, float df)
{
  const float rq = __builtin_amdgcn_rcpf(qf);
  const float d = df * rq, rr = pf * rq;
  const float v = df * __builtin_amdgcn_rcpf(pf + qf);
  const float v2 = v * v;
  float s = 1.f / 13.f;
  s = s * v2 + 1.f / 11.f; s = s * v2 + 1.f / 9.f; s = s * v2 + 1.f / 7.f; s = s * v2 + 1.f / 5.f;
  s = s * v2 + 1.f / 3.f;
  const float ser = d * v + 2.f * rr * (v * v2) * s;
  const float dir = rr * (__builtin_amdgcn_logf(rr) * 0.693147180559945309f) - (rr - 1.f);
  return (fabsf(v) < 0.25f) ? ser : dir;
}
// p ln p for p > 0 (FP32, hardware logarithm)
__device__ __forceinline__ float plogp(float p) { return p * (__builtin_amdgcn_logf(p) * 0.693147180559945309f); }

// FP32 statistics (STAT_FP32): the sums that cancel are formed in FP64, the transcendental and
// nonlinear parts in FP32.  Per cell k: pp_k and q_k = pm_i[x] pm_j[y] and delta_k = pp_k - q_k in
// FP64; then
//   pp ln(pp / q) = q phi(pp / q) + delta,  phi(r) = r ln r - r + 1 >= 0,
// so GT / MI / MIg / MIr become sum_k q_k phi_k (FP32, every term >= 0: no cancellation) plus
// sum_k delta_k (FP64, ~0 for the full 4 x 4 table) -- the same quantity as the reference's
// sum pp (ln pp - ln pm_i - ln pm_j), without the cancellation of its large terms.  CHI and OMES
// are sums of squares of delta (FP32).  The reference's cell conditions are kept (q > 0,
// pp > 0; C2's pooled cells; CWC's allowed cells).  CCF has no logarithm: FP64.
template <int TYPE, int CLS>
__device__ double pair_stat_f32(const StatArgs &A, const double *cd, int64_t i, int64_t j)
{
  double ns = 0.;
#pragma unroll
  for (int k = 0; k < 16; k++) ns += cd[k];
  const int type = TYPE >= 0 ? TYPE : A.type;
  const int cls = CLS >= 0 ? CLS : A.cls;
  double pmi[4], pmj[4];
#pragma unroll
  for (int x = 0; x < 4; x++) { pmi[x] = A.pm[i * 4 + x]; pmj[x] = A.pm[j * 4 + x]; }
  if (type == RSG_CCF) {
    double sa = 0., sb = 0.;
#pragma unroll
    for (int x = 0; x < 4; x++) {
      const double a = ns * pmi[x] - A.meanp[x], b = ns * pmj[x] - A.meanp[x];
      sa += a * a; sb += b * b;
    }
    return sqrt(sa * sb);
  }
  const double rden = 1. / (ns + 16e-10);
  const bool logs = (type == RSG_GT || type == RSG_MI || type == RSG_MIg || type == RSG_MIr);
  // one "cell" of the sums: probability p, expectation q (both FP64)
  float sphi = 0.f, ssq = 0.f, sh = 0.f;
  double sdel = 0.;
  auto cell = [&](double p, double q) {
    const double dl = p - q;
    const float qf = (float)q, df = (float)dl;
    if (logs) {
      if (q > 0. && p > 0.) {
        sphi += qf * phi_pq((float)p, qf, df);
        sdel += dl;
      } else if (p > 0. && cls == RSG_C2 && (type == RSG_MI || type == RSG_MIg)) {
        sphi += INFINITY;                  // C2 MI / MIg take ln q unchecked: q = 0 gives +inf
      }
      if (type == RSG_MIr && p > 0.) sh -= plogp((float)p);
    } else if (q > 0.) {
      ssq += (type == RSG_CHI) ? df * df * __builtin_amdgcn_rcpf(qf) : df * df;
    }
  };
  if (cls == RSG_C2) {
    double pw = 0., pn = 0., qw = 0., qn = 0.;
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int y = 0; y < 4; y++) {
        const double p = (cd[x * 4 + y] + 1e-10) * rden, q = pmi[x] * pmj[y];
        if ((A.allow >> (x * 4 + y)) & 1u) { pw += p; qw += q; } else { pn += p; qn += q; }
      }
    if (type == RSG_CHI || type == RSG_OMES) {
      if (!(ns > 0.)) return 0.;
      cell(pw, qw); cell(pn, qn);
      return ns * (double)ssq;
    }
    if (type == RSG_GT && !(ns > 0.)) return 0.;
    cell(pw, qw); cell(pn, qn);
  } else {
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int y = 0; y < 4; y++) {
        if (cls == RSG_CWC && !((A.allow >> (x * 4 + y)) & 1u)) continue;
        cell((cd[x * 4 + y] + 1e-10) * rden, pmi[x] * pmj[y]);
      }
    if (type == RSG_CHI || type == RSG_OMES) return (ns > 0.) ? ns * (double)ssq : 0.;
  }
  const double s = (double)sphi + sdel;                    // sum pp ln(pp / q)
  switch (type) {
  case RSG_GT:  return (ns > 0.) ? 2. * ns * s : 0.;
  case RSG_MI:  return s;
  case RSG_MIg: {
    // ngap = total - nseff cancels most of nseff: take nseff exact where the counts provide it
    // (the probabilities keep the cells' own sum, which makes sum p = sum q = 1)
    const double ng = A.nsx ? A.nsx[rsg_rec(i, j, A.L)] : ns;
    return s - ((ng > 0.) ? (A.wsum - ng) / ng : 0.);
  }
  default:      return (sh > 1e-2f) ? s / (double)sh : 0.;  // MIr
  }
}

// ---- FP32 statistics of class C16 (GT, MI, MIg, MIr, CHI, OMES): the short form.
// Every cell of the 4 x 4 table has q = pm_i[x] pm_j[y] > 0 (the marginals carry the 1e-10
// pseudo-counts of every pair, or are uniform).  With c' = c + 1e-10, den = ns + 16e-10 and the
// expected count B = den q, the cell's term of sum pp ln(pp / q) splits as
//   pp ln(pp / q) = F / den + (pp - q),   F = c' ln(c' / B) - (c' - B) >= 0,
// and the second parts sum to 1 - S_i S_j, S = sum_x pm[x] (FP64, ~1e-16).  B and e = c - B
// are formed in FP64 (full-rate FMA units on this chip), so e carries no cancellation; the rest
// is FP32.  With v = e / (c + B) = e / (2c - e) (C. Loader, "Fast and accurate computation of
// binomial probabilities", 2000):
//   |v| < 0.3:  F = v (e + c u S(u)),  u = v^2,  c u S(u) = 2 c sum_{k=1..6} v^(2k) / (2k + 1)
//               (truncation < 2e-8 of F)
//   otherwise:  F = c' ln(c' / B) - e   (c' / B with one Newton step, hardware logarithm:
//               error below 6e-7 of F)
// c' and c differ by 1e-10, which changes e and c + B by less than 1e-10 of any F that counts;
// a zero count has v = -1 and takes the second form with c' = 1e-10.  The logarithm is
// evaluated only in waves where some lane needs it (a wave-uniform branch; about 2% of the
// cells of a wave at N = 5000, L = 1500).  MIr divides by H = -sum pp ln pp (hardware logarithm
// per cell, as pair_stat_f32); CHI and OMES are (ns / den) sum e^2 / B and (ns / den^2) sum e^2.
// c: the 16 counts as FP32 (exact for uint16 records, the record values for FP32 ones);
// nsd: their FP64 sum; pi: pm_i; si: sum_x pm_i[x].
template <int KIND, int TYPE>
__device__ __forceinline__ double pair_stat_f32c16(const StatArgs &A, const float *c, double nsd,
                                                   const double *pi, double si, int64_t i, int64_t j)
{
  const double den = nsd + 16e-10;
  double pj[4];
#pragma unroll
  for (int y = 0; y < 4; y++) pj[y] = A.pm[j * 4 + y];
  if constexpr (TYPE == RSG_CHI || TYPE == RSG_OMES) {
    // ns sum (pp - q)^2 / q = (ns / den) sum e^2 / B;  ns sum (pp - q)^2 = (ns / den^2) sum e^2
    float ssq = 0.f;
#pragma unroll
    for (int x = 0; x < 4; x++) {
      const double ax = den * pi[x];
#pragma unroll
      for (int y = 0; y < 4; y++) {
        const double B = ax * pj[y];
        const float e = (float)((double)c[x * 4 + y] - B);
        ssq = (TYPE == RSG_CHI) ? fmaf(e * e, __builtin_amdgcn_rcpf((float)B), ssq) : fmaf(e, e, ssq);
      }
    }
    if (!(nsd > 0.)) return 0.;
    return (TYPE == RSG_CHI) ? nsd / den * (double)ssq : nsd / (den * den) * (double)ssq;
  }
  const double sj = (pj[0] + pj[1]) + (pj[2] + pj[3]);
  const float rdf = __builtin_amdgcn_rcpf((float)den);
  float hsum = 0.f;                                      // MIr: sum pp log2 pp
  // the series coefficients 2 / (2k + 1) as register operands (empty asm statements pin them:
  // 2 / 13 in a VGPR, the rest in SGPRs, one per instruction), so no cell spends vector moves
  // re-creating them
  float k3 = 2.f / 3.f, k5 = 2.f / 5.f, k7 = 2.f / 7.f, k9 = 2.f / 9.f, k11 = 2.f / 11.f, k13 = 2.f / 13.f;
  asm("" : "+s"(k3), "+s"(k5), "+s"(k7), "+s"(k9), "+s"(k11));
  asm("" : "+v"(k13));
  float acc = 0.f;
#pragma unroll
  for (int x = 0; x < 4; x++) {
    const double ax = den * pi[x];
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const float cf = c[x * 4 + y];
      const double B = ax * pj[y];
      const float e = (float)((double)cf - B);
      const float v = e * __builtin_amdgcn_rcpf(fmaf(2.f, cf, -e));
      const float u = v * v;
      float s = fmaf(k13, u, k11);
      s = fmaf(s, u, k9); s = fmaf(s, u, k7); s = fmaf(s, u, k5); s = fmaf(s, u, k3);
      float F = v * fmaf(cf * u, s, e);
      const bool far = fabsf(v) >= 0.3f;
      if (__builtin_amdgcn_ballot_w64(far)) {
        const float bh = (float)B, ch = fmaxf(cf, 1e-10f);
        const float rb = __builtin_amdgcn_rcpf(bh);
        float r = ch * rb;
        r = fmaf(fmaf(-bh, r, ch), rb, r);
        const float Fl = fmaf(ch * 0.693147180559945309f, __builtin_amdgcn_logf(r), -e);
        F = far ? Fl : F;
      }
      acc += F;
      if constexpr (TYPE == RSG_MIr) {
        const float pp = fmaxf(cf, 1e-10f) * rdf;
        hsum = fmaf(pp, __builtin_amdgcn_logf(pp), hsum);
      }
    }
  }
  const double s = (double)(acc * rdf) + fma(-si, sj, 1.0);   // sum pp ln(pp / q)
  if constexpr (TYPE == RSG_GT) return (nsd > 0.) ? 2. * nsd * s : 0.;
  if constexpr (TYPE == RSG_MI) return s;
  if constexpr (TYPE == RSG_MIr) {                                 // MI / H, H = -sum pp ln pp
    const float H = -hsum * 0.693147180559945309f;
    return (H > 1e-2f) ? s / (double)H : 0.;
  }
  const double ng = A.nsx ? A.nsx[rsg_rec(i, j, A.L)] : nsd;        // MIg
  return s - ((ng > 0.) ? (A.wsum - ng) / ng : 0.);
}

// A record as 16 FP32 counts and their FP64 sum (integer sum for uint16 records).
template <int KIND> __device__ __forceinline__ double decode_rec_f32(const RawRec<KIND> &r, float *c)
{
  if constexpr (KIND == CNT_U16) {
    unsigned ns = 0;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const unsigned w[4] = {r.q[h].x, r.q[h].y, r.q[h].z, r.q[h].w};
#pragma unroll
      for (int q = 0; q < 4; q++) {
        c[8 * h + 2 * q] = (float)(w[q] & 0xFFFFu); c[8 * h + 2 * q + 1] = (float)(w[q] >> 16);
        ns += (w[q] & 0xFFFFu) + (w[q] >> 16);
      }
    }
    return (double)ns;
  } else {
    double ns = 0.;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint4 v = r.q[k];
      c[4 * k] = __uint_as_float(v.x); c[4 * k + 1] = __uint_as_float(v.y);
      c[4 * k + 2] = __uint_as_float(v.z); c[4 * k + 3] = __uint_as_float(v.w);
    }
#pragma unroll
    for (int k = 0; k < 16; k++) ns += (double)c[k];
    return ns;
  }
}

// One pair. c: counts, pmi/pmj: marginals.  Returns the statistic (see the file header).
// TYPE / CLS >= 0 fix the statistic and class at compile time (-1: read them from A).
// tabc: ln(c + 1e-10) for counts c = 0..N (A.lgc, or a copy of it in LDS) when TAB; A.lgc holds
// ln(c + 1e-10), ln(c + 16e-10) and (c + 1e-10) ln(c + 1e-10) for c = 0..N in three runs.
template <typename T, bool TAB, int TYPE = -1, int CLS = -1>
__device__ double pair_stat(const StatArgs &A, const double *cd, int64_t i, int64_t j, const double *tabc)
{
#pragma clang fp contract(off)   // no FMA: every specialisation and the batched path round alike
  if constexpr (sizeof(T) == sizeof(float)) return pair_stat_f32<TYPE, CLS>(A, cd, i, j);
  T ns = 0;
  double nsd = 0.;
#pragma unroll
  for (int k = 0; k < 16; k++) { ns += (T)cd[k]; nsd += cd[k]; }
  const double nsg = (!TAB && A.nsx) ? A.nsx[rsg_rec(i, j, A.L)] : nsd;   // MIg's nseff (exact where given)
  T pmi[4], pmj[4], lpi[4], lpj[4];
#pragma unroll
  for (int x = 0; x < 4; x++) {
    pmi[x] = (T)A.pm[i * 4 + x]; pmj[x] = (T)A.pm[j * 4 + x];
    lpi[x] = (T)A.lpm[i * 4 + x]; lpj[x] = (T)A.lpm[j * 4 + x];
  }
  const T den = ns + (T)16e-10;
  T lden;
  if (TAB) lden = (T)A.lgc[A.N + 1 + (int64_t)nsd]; else lden = tlog<T>(den);
  // pp and ln pp of cell k, formed where they are used (no 16-entry arrays kept live); one
  // reciprocal of the normaliser instead of 16 divisions (pp within 1 ulp of the quotient)
  const T rden = (T)1 / den;
  auto PP = [&](int k) -> T { return ((T)cd[k] + (T)1e-10) * rden; };
  auto LPP = [&](int k) -> T {
    if (TAB) return (T)tabc[(int)cd[k]] - lden;
    return tlog<T>((T)cd[k] + (T)1e-10) - lden;
  };
  const int type = TYPE >= 0 ? TYPE : A.type;
  const int cls = CLS >= 0 ? CLS : A.cls;
  if (type == RSG_CCF) {
    T sa = 0, sb = 0;
#pragma unroll
    for (int x = 0; x < 4; x++) {
      T a = ns * pmi[x] - (T)A.meanp[x], b = ns * pmj[x] - (T)A.meanp[x];
      sa += a * a; sb += b * b;
    }
    return (double)sqrt((double)(sa * sb));
  }
  if (cls == RSG_C2) {
    T pw = 0, pn = 0, qw = 0, qn = 0;
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int y = 0; y < 4; y++) {
        if ((A.allow >> (x * 4 + y)) & 1u) { pw += PP(x * 4 + y); qw += pmi[x] * pmj[y]; }
        else                               { pn += PP(x * 4 + y); qn += pmi[x] * pmj[y]; }
      }
    T v = 0, H = 0;
    switch (type) {
    case RSG_CHI:
      if (ns > 0) { if (qw > 0) v += ns * (pw - qw) * (pw - qw) / qw; if (qn > 0) v += ns * (pn - qn) * (pn - qn) / qn; }
      break;
    case RSG_OMES:
      if (ns > 0) { if (qw > 0) v += ns * (pw - qw) * (pw - qw); if (qn > 0) v += ns * (pn - qn) * (pn - qn); }
      break;
    case RSG_GT:
      if (ns > 0) {
        if (qw > 0 && pw > 0) v += pw * (tlog<T>(pw) - tlog<T>(qw));
        if (qn > 0 && pn > 0) v += pn * (tlog<T>(pn) - tlog<T>(qn));
        v *= 2 * ns;
      }
      break;
    case RSG_MI: case RSG_MIg:
      if (pw > 0) v += pw * (tlog<T>(pw) - tlog<T>(qw));
      if (pn > 0) v += pn * (tlog<T>(pn) - tlog<T>(qn));
      if (type == RSG_MIg && ns > 0) v -= (T)((A.wsum - nsg) / nsg);
      break;
    case RSG_MIr:
      if (pw > 0) H -= pw * tlog<T>(pw);
      if (pn > 0) H -= pn * tlog<T>(pn);
      if (pw > 0 && qw > 0) v += pw * (tlog<T>(pw) - tlog<T>(qw));
      if (pn > 0 && qn > 0) v += pn * (tlog<T>(pn) - tlog<T>(qn));
      v = (H > (T)1e-2) ? v / H : 0;
      break;
    }
    return (double)v;
  }
  // C16 / CWC
  if (TAB && cls == RSG_C16 && (type == RSG_GT || type == RSG_MI || type == RSG_MIr || type == RSG_MIg)) {
    // GT and the MI family over all 16 cells, regrouped: pm_i[x] pm_j[y] > 0 for every cell (the
    // marginals carry the 1e-10 pseudo-counts of every pair, or are uniform), so
    //   sum_xy pp (ln pp - ln pm_i[x] - ln pm_j[y]) = rden sum_xy (c + 1e-10) ln(c + 1e-10)
    //     - ln(den) rden (ns + 16e-10) - rden sum_x ln pm_i[x] (r_x + 4e-10)
    //     - rden sum_y ln pm_j[y] (s_y + 4e-10)
    // with r_x, s_y the row and column sums of the counts
    T a = 0, r[4] = {0, 0, 0, 0}, cs[4] = {0, 0, 0, 0};
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int y = 0; y < 4; y++) {
        const int k = x * 4 + y;
        a += (T)A.lgc[2 * (A.N + 1) + (int64_t)cd[k]];      // (c + 1e-10) ln(c + 1e-10)
        r[x] += (T)cd[k]; cs[y] += (T)cd[k];
      }
    T b = 0;
#pragma unroll
    for (int x = 0; x < 4; x++) b += lpi[x] * (r[x] + (T)4e-10) + lpj[x] * (cs[x] + (T)4e-10);
    const T s = rden * (a - b) - lden * (rden * den);               // MI
    switch (type) {
    case RSG_GT:  return (ns > 0) ? 2. * (double)ns * (double)s : 0.;
    case RSG_MI:  return (double)s;
    case RSG_MIg: return (double)s - ((nsd > 0) ? (A.wsum - nsd) / nsd : 0.);
    default: {                                                       // MIr: MI / H, H = -sum pp ln pp
      const T H = -(rden * a - lden * (rden * den));
      return (H > (T)1e-2) ? (double)(s / H) : 0.;
    }
    }
  }
  T s = 0, H = 0;
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const int k = x * 4 + y;
      const T q = pmi[x] * pmj[y];
      switch (type) {
      case RSG_CHI:  if (q > 0) { const T d = PP(k) - q; s += d * d / q; } break;
      case RSG_OMES: if (q > 0) { const T d = PP(k) - q; s += d * d; } break;
      case RSG_GT:
        if (cls == RSG_CWC && !((A.allow >> k) & 1u)) break;
        if (q > 0) s += PP(k) * (LPP(k) - lpi[x] - lpj[y]);
        break;
      case RSG_MI: case RSG_MIg: case RSG_MIr: {
        const T pp = PP(k), lpp = LPP(k);
        if (pmi[x] > 0 && pmj[y] > 0) s += pp * (lpp - lpi[x] - lpj[y]);
        if (type == RSG_MIr) H -= pp * lpp;
        break;
      }
      }
    }
  switch (type) {
  case RSG_CHI: case RSG_OMES: return (ns > 0) ? (double)(ns * s) : 0.;
  case RSG_GT:   return (ns > 0) ? 2. * (double)ns * (double)s : 0.;
  case RSG_MI:   return (double)s;
  case RSG_MIg:  return (double)s - ((nsg > 0) ? (A.wsum - nsg) / nsg : 0.);
  case RSG_MIr:  return (H > (T)1e-2) ? (double)(s / H) : 0.;
  }
  return 0.;
}

// ---- corr_Marginals, pm[i][x] = normalize_x(sum over j != i with nseff_ij > 0 of the pp_ij
// row (i first) or column (i second) sums).  Pass 1 (pm_tile, 64 * PMW threads) reads the
// records of 64 i-columns x PMJ j-columns (sub-block `sub` of a 64 x 64 tile) once and
// leaves per-block partial sums for the 64 i-columns (row sums) and per-tile sums for the PMJ
// j-columns (column sums); pass 2 (pm_reduce_col, one wave per column) adds a column's
// partials in a fixed order (deterministic) and normalises.  The single-alignment and the
// batched paths share both, so they round identically.
#define PMT 64
#define PMS 8           // j sub-blocks per 64 x 64 tile (one work group each)
#define PMJ (PMT / PMS) // j-columns per work group
#define PMW 4           // waves per work group: PMJ / PMW = 2 j-columns per thread
// partials of upper-triangle tile (bi <= bj) of nbt x nbt tiles, row-major
__host__ __device__ __forceinline__ int64_t pm_tri(int64_t bi, int64_t bj, int64_t nbt)
{
  return bi * nbt - (bi * (bi - 1)) / 2 + (bj - bi);
}
// doubles of the i-side partials [tri][PMS][64][4] and of the j-side partials [tri][64][4]
__host__ __device__ __forceinline__ int64_t pm_npi(int64_t nbt) { return pm_tri(nbt, nbt, nbt) * PMS * PMT * 4; }
__host__ __device__ __forceinline__ int64_t pm_npj(int64_t nbt) { return pm_tri(nbt, nbt, nbt) * PMT * 4; }
template <int KIND>
__device__ void pm_tile(const void *__restrict__ cnt, int64_t L, int nbt, int bi, int bj, int sub,
                        double *__restrict__ part_i, double *__restrict__ part_j)
{
  __shared__ double si[PMW][PMT][4], sj[PMJ][4];
  const int tid = threadIdx.x, ii = tid & 63, jq = tid >> 6;
  const int64_t i = (int64_t)bi * PMT + ii;
  double ai[4] = {0., 0., 0., 0.};
#pragma unroll
  for (int q = 0; q < PMJ / PMW; q++) {
    const int js = jq + q * PMW, jj = sub * PMJ + js;       // js: column within the sub-block
    const int64_t j = (int64_t)bj * PMT + jj;
    double cj[4] = {0., 0., 0., 0.};
    if (i < j && j < L) {
      double cd[16];
      load_rec<KIND>(cnt, rsg_rec(i, j, L), cd);
      double ns = 0.;
#pragma unroll
      for (int k = 0; k < 16; k++) ns += cd[k];
      if (ns > 0) {
        const double inv = 1. / (ns + 16e-10);
#pragma unroll
        for (int x = 0; x < 4; x++) {
          ai[x] += (cd[x * 4] + cd[x * 4 + 1] + cd[x * 4 + 2] + cd[x * 4 + 3] + 4e-10) * inv;
          cj[x] = (cd[x] + cd[4 + x] + cd[8 + x] + cd[12 + x] + 4e-10) * inv;
        }
      }
    }
#pragma unroll
    for (int x = 0; x < 4; x++)
      for (int o = 32; o > 0; o >>= 1) cj[x] += __shfl_down(cj[x], o, 64);
    if (ii == 0) for (int x = 0; x < 4; x++) sj[js][x] = cj[x];
  }
  for (int x = 0; x < 4; x++) si[jq][ii][x] = ai[x];
  __syncthreads();
  const int64_t t = pm_tri(bi, bj, nbt);
  const int c = tid >> 2, x = tid & 3;                       // 256 threads: 64 columns x 4
  double a = 0.;
  for (int w = 0; w < PMW; w++) a += si[w][c][x];
  part_i[((t * PMS + sub) * PMT + c) * 4 + x] = a;
  if (c < PMJ) part_j[(t * PMT + sub * PMJ + c) * 4 + x] = sj[c][x];
}

// ---- corr_Marginals from unit-weight counts, fused into the int8 Gram.  The partial sums are
// formed per 32 x 32 column block (bi <= bj, in 32-column units) in the MFMA result layout:
// lane (col = lane & 31, khalf = lane >> 5) holds i = 32 bi + col and, for r = 0..15,
// j = 32 bj + 8 (r / 4) + 4 khalf + r % 4.  Per pair: inv = 1 / (ns + 16e-10), row term
// (rowsum_x + 4e-10) inv, column term (colsum_y + 4e-10) inv, with ns, rowsum and colsum exact
// integers; pairs with i >= j, j >= L or ns = 0 have inv = 0 and add +0.
//   i side: a lane adds its row terms over r = 0..7 and over r = 8..15, each in order, then the
//   two sums, then the two khalf lanes of a column (pm_khalf_sum).
//   j side: the 32 column terms of r = 8 h .. 8 h + 7 (index 4 (r - 8 h) + y) are summed over the
//   32 lanes of a khalf by pm_rscatter32, after which lane col holds index col.
// The Gram epilogue, the split-K fix-up and the batched Gram all follow this, so the partials do
// not depend on tile shape or K splitting.  Layout: part_i[blk][32][4] and part_j[blk][32][4],
// blk = pm_tri(bi, bj, nb32).
__host__ __device__ __forceinline__ int64_t pm32_n(int64_t nb32) { return pm_tri(nb32, nb32, nb32) * 32 * 4; }

__device__ __forceinline__ double pm_inv(bool valid, int ns)
{
#pragma clang fp contract(off)
  return (valid && ns > 0) ? 1. / ((double)ns + 16e-10) : 0.;
}
__device__ __forceinline__ double pm_term(int sum, double inv)
{
#pragma clang fp contract(off)
  return ((double)sum + 4e-10) * inv;
}
// i side: the khalf halves of a lane pair combined (a + b is exact-commutative, so both lanes agree)
__device__ __forceinline__ double pm_khalf_sum(double a) { return a + __shfl_xor(a, 32, 64); }
// j side: reduce-scatter over the 32 lanes of a khalf.  Butterfly over lane bits 16, 8, 4, 2, 1:
// at each step a lane keeps the lower half of its values if its bit is clear, the upper half
// otherwise, and adds the partner's copy of the half it keeps.  Lane col ends with the sum over
// the 32 lanes of value index col (31 exchanges instead of 32 x 5).
// (The stage width is a template argument so that every index is a constant and v stays in
// registers.)
template <int M>
__device__ __forceinline__ void pm_rscatter_step(double (&v)[32], int col)
{
  const bool up = (col & M) != 0;
#pragma unroll
  for (int k = 0; k < M; k++) {
    const double send = up ? v[k] : v[k + M];
    const double keep = up ? v[k + M] : v[k];
    v[k] = keep + __shfl_xor(send, M, 64);
  }
}
__device__ __forceinline__ double pm_rscatter32(double (&v)[32], int col)
{
  pm_rscatter_step<16>(v, col);
  pm_rscatter_step<8>(v, col);
  pm_rscatter_step<4>(v, col);
  pm_rscatter_step<2>(v, col);
  pm_rscatter_step<1>(v, col);
  return v[0];
}

// 16 lanes (t = 0..15) per column: lane t adds blocks k = t, t + 16, ... of the column's list
// (blocks (b, b..nb32-1) on the i side, then (0..b, b) on the j side) in order, an xor tree over
// the 16 lanes combines them (every lane ends with the same totals) and lane x < 4 writes
// pm[col][x] and its log.  Every block holds a pair i < j < L, so its tile is computed, except a
// last diagonal block with a single column (32 b + 1 >= L), which no producer writes and which
// is skipped here.
__device__ __forceinline__ void pm32_reduce_col(const double *__restrict__ part_i, const double *__restrict__ part_j,
                                                int nb32, int64_t L, int64_t col, int t, double tol,
                                                double *__restrict__ pm, double *__restrict__ lpm, int *__restrict__ flag)
{
  const int b = (int)(col / 32), o = (int)(col % 32);
  const bool diag = 32 * (int64_t)b + 1 < L;
  double v[4] = {0., 0., 0., 0.};
  for (int k = t; k < nb32 + 1; k += 16) {
    int64_t base;
    const double *src;
    if (k < nb32 - b) { if (k == 0 && !diag) continue; src = part_i; base = pm_tri(b, b + k, nb32) * 32 + o; }
    else if (k - (nb32 - b) <= b) { if (k - (nb32 - b) == b && !diag) continue; src = part_j; base = pm_tri(k - (nb32 - b), b, nb32) * 32 + o; }
    else continue;
    const double4 q = *(const double4 *)(src + base * 4);
    v[0] += q.x; v[1] += q.y; v[2] += q.z; v[3] += q.w;
  }
#pragma unroll
  for (int off = 8; off > 0; off >>= 1)
#pragma unroll
    for (int x = 0; x < 4; x++) v[x] += __shfl_xor(v[x], off, 16);
  if (t >= 4) return;
  const double sum = v[0] + v[1] + v[2] + v[3];
  double p[4], tot = 0.;
#pragma unroll
  for (int x = 0; x < 4; x++) { p[x] = (sum != 0.) ? v[x] / sum : 0.25; tot += p[x]; }
  const double px = (t == 0) ? p[0] : (t == 1) ? p[1] : (t == 2) ? p[2] : p[3];
  if (!(px >= 0. && px <= 1.) || (t == 0 && fabs(tot - 1.) > tol)) atomicOr(flag, 1);
  pm[col * 4 + t] = px;
  lpm[col * 4 + t] = (px > 0.) ? log(px) : -INFINITY;
}

// one wave (lane = 0..63) per column: lanes take the column's tiles in a fixed assignment,
// a fixed shuffle tree adds them
__device__ __forceinline__ void pm_reduce_col(const double *__restrict__ part_i, const double *__restrict__ part_j,
                                              int nbt, int64_t L, int64_t col, int lane, double tol,
                                              double *__restrict__ pm, double *__restrict__ lpm, int *__restrict__ flag)
{
  const int b = (int)(col / PMT), o = (int)(col % PMT);
  const int ni = PMS * (nbt - b);                             // i-side entries: tiles (b, b..nbt-1) x sub-blocks
  double v[4] = {0., 0., 0., 0.};
  for (int k = lane; k < ni + b + 1; k += 64) {
    // k < ni: i-side partial of tile (b, b + k / PMS), sub-block k % PMS; else j-side partial of
    // tile (k - ni, b)
    int64_t base;
    const double *src;
    if (k < ni) { src = part_i; base = (pm_tri(b, b + k / PMS, nbt) * PMS + k % PMS) * PMT + o; }
    else { src = part_j; base = pm_tri(k - ni, b, nbt) * PMT + o; }
    for (int x = 0; x < 4; x++) v[x] += src[base * 4 + x];
  }
#pragma unroll
  for (int x = 0; x < 4; x++)
    for (int off = 32; off > 0; off >>= 1) v[x] += __shfl_down(v[x], off, 64);
  if (lane != 0) return;
  const double sum = v[0] + v[1] + v[2] + v[3];
  double tot = 0.;
  for (int x = 0; x < 4; x++) {
    v[x] = (sum != 0.) ? v[x] / sum : 0.25;
    if (!(v[x] >= 0. && v[x] <= 1.)) atomicOr(flag, 1);
    tot += v[x];
    pm[col * 4 + x] = v[x];
    lpm[col * 4 + x] = (v[x] > 0.) ? log(v[x]) : -INFINITY;
  }
  if (fabs(tot - 1.) > tol) atomicOr(flag, 1);
}

// 32 x 32 pair tile (bi <= bj); COV written to both triangles through an LDS transpose;
// the thread's min / max folded into lmin / lmax.
template <typename T, int KIND, bool TAB, int TYPE = -1, int CLS = -1>
__device__ __forceinline__ void stat_pairs(const StatArgs &A, int bi, int bj, const double *tabc, double &lmin, double &lmax)
{
  __shared__ double tile[32][33];
  const int tid = threadIdx.x;
  {
    // phase 1: consecutive threads take consecutive i (consecutive records j*L + i),
    // write row j of COV; phase 2 writes row i from the LDS transpose.  The four records of
    // a thread are loaded before any statistic is formed and every COV store waits until all
    // four are done, so the loads (and the table reads they feed) of the four pairs overlap.
    const int ii = tid & 31;
    const int64_t i = (int64_t)bi * 32 + ii;
    double v[4];
    if constexpr (KIND == CNT_F32) {
      // 64-byte (weighted) records: with two records in flight (the next one loads while this
      // one is scored) the FP32 statistic needs ~100 VGPRs and runs four waves per SIMD; four
      // records in flight took 176 VGPRs and two waves
      RawRec<KIND> cur, nxt;
      {
        const int64_t j = (int64_t)bj * 32 + (tid >> 5);
        if (i < j && j < A.L) load_raw<KIND>(A.cnt, rsg_rec(i, j, A.L), cur);
      }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int64_t j = (int64_t)bj * 32 + (tid >> 5) + 8 * k;
        if (k < 3) {
          const int64_t j2 = j + 8;
          if (i < j2 && j2 < A.L) load_raw<KIND>(A.cnt, rsg_rec(i, j2, A.L), nxt);
        }
        v[k] = -INFINITY;
        if (i < j && j < A.L) {
          double cd[16];
          decode_rec<KIND>(cur, cd);
          v[k] = pair_stat<T, TAB, TYPE, CLS>(A, cd, i, j, tabc);
          lmin = fmin(lmin, v[k]); lmax = fmax(lmax, v[k]);
        }
        if (k < 3) cur = nxt;
      }
    } else {
      RawRec<KIND> raw[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int64_t j = (int64_t)bj * 32 + (tid >> 5) + 8 * k;
        if (i < j && j < A.L) load_raw<KIND>(A.cnt, rsg_rec(i, j, A.L), raw[k]);
      }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int64_t j = (int64_t)bj * 32 + (tid >> 5) + 8 * k;
        v[k] = -INFINITY;
        if (i < j && j < A.L) {
          double cd[16];
          decode_rec<KIND>(raw[k], cd);
          v[k] = pair_stat<T, TAB, TYPE, CLS>(A, cd, i, j, tabc);
          lmin = fmin(lmin, v[k]); lmax = fmax(lmax, v[k]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int jj = (tid >> 5) + 8 * k;
      const int64_t j = (int64_t)bj * 32 + jj;
      if (i < j && j < A.L) A.cov[j * A.L + i] = v[k];
      else if (i == j && i < A.L) A.cov[i * A.L + j] = -INFINITY;
      tile[jj][ii] = v[k];
    }
    __syncthreads();
    const int jj = tid & 31;
    for (int k = 0; k < 4; k++) {
      const int ii2 = (tid >> 5) + 8 * k;
      const int64_t i = (int64_t)bi * 32 + ii2, j = (int64_t)bj * 32 + jj;
      if (i < j && j < A.L) A.cov[i * A.L + j] = tile[jj][ii2];
    }
    __syncthreads();                      // the tile buffer is reused by the next call
  }
}

// FP32 short-form statistics (class C16: pair_stat_f32c16) over half of a 32 x 32
// pair tile (bi <= bj): columns j = bj*32 + 16h .. + 15 against i = bi*32 .. + 31, two pairs per
// thread, so a grid has twice the blocks of whole tiles and its last blocks hold less work.
// COV is written to both triangles (row j directly, row i through an LDS transpose); the
// thread's min / max are folded into lmin / lmax.
template <int KIND, int TYPE>
__device__ __forceinline__ void stat_half_f32c16(const StatArgs &A, int bi, int bj, int h, double &lmin, double &lmax)
{
  __shared__ double tile[16][33];
  const int tid = threadIdx.x, ii = tid & 31;
  const int64_t i = (int64_t)bi * 32 + ii, jb = (int64_t)bj * 32 + 16 * h;
  double pi[4] = {0., 0., 0., 0.}, si = 0.;
  if (i < A.L) {
#pragma unroll
    for (int x = 0; x < 4; x++) pi[x] = A.pm[i * 4 + x];
    si = (pi[0] + pi[1]) + (pi[2] + pi[3]);
  }
  RawRec<KIND> raw[2];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int64_t j = jb + (tid >> 5) + 8 * k;
    if (i < j && j < A.L) load_raw<KIND>(A.cnt, rsg_rec(i, j, A.L), raw[k]);
  }
  double v[2];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int64_t j = jb + (tid >> 5) + 8 * k;
    v[k] = -INFINITY;
    if (i < j && j < A.L) {
      float cf[16];
      const double nsd = decode_rec_f32<KIND>(raw[k], cf);
      v[k] = pair_stat_f32c16<KIND, TYPE>(A, cf, nsd, pi, si, i, j);
      lmin = fmin(lmin, v[k]); lmax = fmax(lmax, v[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int jl = (tid >> 5) + 8 * k;
    const int64_t j = jb + jl;
    if (i < j && j < A.L) A.cov[j * A.L + i] = v[k];
    else if (i == j && i < A.L) A.cov[i * A.L + j] = -INFINITY;
    tile[jl][ii] = v[k];
  }
  __syncthreads();
  const int jl = tid & 15;
  const int64_t j = jb + jl;
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int ii2 = (tid >> 4) + 16 * k;
    const int64_t i2 = (int64_t)bi * 32 + ii2;
    if (i2 < j && j < A.L) A.cov[i2 * A.L + j] = tile[jl][ii2];
  }
}

// Sum of v over a 256-thread block with the pairing of the LDS tree "for st = 128 .. 1:
// red[t] += red[t + st]": the first two levels in LDS, the last six by wave
// shuffles in wave 0 -- the same additions in the same order, so the same value, with three
// barriers instead of eight.  Every thread gets the sum (sh: 257 doubles, sh[256] the result).
__device__ __forceinline__ double block_sum_tree256(double v, double *sh)
{
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  if (t < 128) sh[t] += sh[t + 128];
  __syncthreads();
  if (t < 64) {
    double x = sh[t] + sh[t + 64];
#pragma unroll
    for (int st = 32; st > 0; st >>= 1) x += __shfl_down(x, st, 64);
    if (t == 0) sh[256] = x;
  }
  __syncthreads();
  return sh[256];
}

// Min and max over a 256-thread block, the same way (sa, sb: 256 doubles each); the result is
// valid in thread 0.
__device__ __forceinline__ void block_minmax256(double &a, double &b, double *sa, double *sb)
{
  const int t = threadIdx.x;
  sa[t] = a; sb[t] = b;
  __syncthreads();
  if (t < 128) { sa[t] = fmin(sa[t], sa[t + 128]); sb[t] = fmax(sb[t], sb[t + 128]); }
  __syncthreads();
  if (t < 64) {
    a = fmin(sa[t], sa[t + 64]); b = fmax(sb[t], sb[t + 64]);
#pragma unroll
    for (int st = 32; st > 0; st >>= 1) { a = fmin(a, __shfl_down(a, st, 64)); b = fmax(b, __shfl_down(b, st, 64)); }
  }
}

// Block min / max of the 256 threads to part[2 * part_idx], part[2 * part_idx + 1].
__device__ __forceinline__ void block_minmax(double lmin, double lmax, double *part, int64_t part_idx)
{
  __shared__ double rmin[256], rmax[256];
  block_minmax256(lmin, lmax, rmin, rmax);
  if (threadIdx.x == 0) { part[2 * part_idx] = lmin; part[2 * part_idx + 1] = lmax; }
}

// Block t of a triangular grid -> pair tile (bi, bj), bi <= bj, column-block major:
// t = bj (bj + 1) / 2 + bi.
__device__ __forceinline__ void tri_tile(int64_t t, int &bi, int &bj)
{
  int64_t b = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((b + 1) * (b + 2) / 2 <= t) b++;
  while (b * (b + 1) / 2 > t) b--;
  bj = (int)b; bi = (int)(t - b * (b + 1) / 2);
}

// One tile per block (tiles below the diagonal contribute +-inf), block min / max to part.
template <typename T, int KIND, bool TAB, int TYPE = -1, int CLS = -1>
__device__ void stat_tile(const StatArgs &A, int bi, int bj, int64_t part_idx)
{
  double lmin = INFINITY, lmax = -INFINITY;
  if (bi <= bj) stat_pairs<T, KIND, TAB, TYPE, CLS>(A, bi, bj, A.lgc, lmin, lmax);
  block_minmax(lmin, lmax, A.part, part_idx);
}

// ---- FP64 null tail fit (cov_histogram_pmass, SetTailByMass, binned exponential / Gamma ML,
// SetSurvFitTail); one work group; out = {mu, lambda, tau, newmass, pmass, phi, cmin, No}.
#include <cfloat>
#include <climits>
__device__ double d_digamma(double x)
{
  double r = 0.;
  while (x < 6.) { r -= 1. / x; x += 1.; }
  double f = 1. / (x * x);
  return r + log(x) - 0.5 / x - f * (1. / 12 - f * (1. / 120 - f * (1. / 252 - f * (1. / 240 - f / 132))));
}

// Regularized upper incomplete gamma (series / Lentz continued fraction).
__device__ double d_gammq(double a, double x)
{
  if (x <= 0.) return 1.0;
  double gln = lgamma(a);
  if (x < a + 1.) {
    double ap = a, sum = 1. / a, del = sum;
    for (int n = 0; n < 1000; n++) { ap += 1.; del *= x / ap; sum += del; if (fabs(del) < fabs(sum) * 1e-16) break; }
    return 1. - sum * exp(-x + a * log(x) - gln);
  }
  double b = x + 1. - a, cc = 1. / DBL_MIN, d = 1. / b, hh = d;
  for (int i = 1; i < 1000; i++) {
    double an = -i * (i - a);
    b += 2.;
    d = an * d + b; if (fabs(d) < DBL_MIN) d = DBL_MIN;
    cc = b + an / cc; if (fabs(cc) < DBL_MIN) cc = DBL_MIN;
    d = 1. / d;
    double del = d * cc; hh *= del;
    if (fabs(del - 1.) < 1e-16) break;
  }
  return exp(-x + a * log(x) - gln) * hh;
}

// Block-wide reductions over 1024 threads in a fixed order (deterministic).
template <typename T, typename Op>
__device__ T block_reduce(T v, T *sh, Op op)
{
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = op(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  T r = sh[0];
  __syncthreads();
  return r;
}

// Inclusive suffix sum over the threads of the block (v of threads tid..T-1), in LDS.
__device__ long long suffix_scan(long long v, long long *sh)
{
  const int T = blockDim.x, tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = 1; off < T; off <<= 1) {
    const long long add = (tid + off < T) ? sh[tid + off] : 0;
    __syncthreads();
    sh[tid] += add;
    __syncthreads();
  }
  const long long r = sh[tid];
  __syncthreads();
  return r;
}

// One work group.  S[b] = sum_{i >= b} obs[i] and NZ[b] = #nonzero bins >= b over [imin, imax]
// (suffix sums), so the reference's downward scans become "largest bin satisfying a monotone
// condition" reductions.
__device__ void nullfit_body(const unsigned long long *obs, int nb, int imin, int imax, double n, double Nc,
                             double bmin, double w, double target_pmass, double fracfit, int doexp,
                             unsigned long long *S, int *NZ, double *survfit, double *out)
{
#pragma clang fp contract(off)   // bin bounds and sums rounded as the host computes them (no FMA)
  __shared__ double shd[1024];
  __shared__ long long shl[1024];
  const int T = blockDim.x, tid = threadIdx.x;
  // suffix sums by chunks: each thread owns a contiguous chunk of bins [imin, imax]
  const int nbin = (imax >= imin) ? imax - imin + 1 : 0;
  const int chunk = (nbin + T - 1) / T;
  const int c0 = imin + tid * chunk, c1 = min(imin + (tid + 1) * chunk, imax + 1);
  unsigned long long csum = 0; long long cnz = 0;
  for (int b = c0; b < c1; b++) { csum += obs[b]; cnz += (obs[b] > 0); }
  // exclusive suffix of the chunk totals (sum over threads with a larger index): an
  // inclusive Hillis-Steele scan from the top, then subtract the thread's own total
  long long above = suffix_scan(csum, shl), abovenz = suffix_scan(cnz, shl);
  above -= (long long)csum; abovenz -= cnz;
  unsigned long long run = (unsigned long long)above; long long runnz = abovenz;
  for (int b = c1 - 1; b >= c0; b--) { run += obs[b]; runnz += (obs[b] > 0); S[b] = run; NZ[b] = (int)runnz; }
  __syncthreads();
  const long long tp = (nbin > 0) ? NZ[imin] : 0;
  // cov_histogram_pmass: the largest nonzero bin where NZ/tp >= fracfit or S/Nc >= target
  long long bsel = LLONG_MIN;
  for (int b = imin + tid; b <= imax; b += T)
    if (obs[b] > 0 && ((double)NZ[b] / (double)tp >= fracfit || (double)S[b] / Nc >= target_pmass))
      bsel = max(bsel, (long long)b);
  bsel = block_reduce<long long>(bsel, shl, [](long long a, long long b) { return a > b ? a : b; });
  double pmass;
  if (bsel == LLONG_MIN) {                       // the scan ran through: pmass of the last nonzero bin
    long long lo = LLONG_MAX;
    for (int b = imin + tid; b <= imax; b += T) if (obs[b] > 0) lo = min(lo, (long long)b);
    lo = block_reduce<long long>(lo, shl, [](long long a, long long b) { return a < b ? a : b; });
    pmass = (lo == LLONG_MAX) ? NAN : (double)S[lo] / Nc;
  } else pmass = (double)S[bsel] / Nc;
  // SetTailByMass: the largest b in [imin, imax] with S[b] >= pmass * n
  long long bt = LLONG_MIN;
  for (int b = imin + tid; b <= imax; b += T) if ((double)S[b] >= pmass * n) bt = max(bt, (long long)b);
  bt = block_reduce<long long>(bt, shl, [](long long a, long long b) { return a > b ? a : b; });
  long long cmin = (bt == LLONG_MIN) ? (long long)imin - 1 : bt;
  unsigned long long kept = (bt == LLONG_MIN) ? ((nbin > 0) ? S[imin] : 0ull) : S[bt];
  const double phi = w * (double)cmin + bmin;
  const double newmass = (double)kept / n;
  // fits over bins [cmin, imax]
  double mu = phi, lambda = INFINITY, tau = 0.;
  if (doexp) {
    double sa = 0., sb = 0.;
    for (int b = (int)max(cmin, (long long)imin) + tid; b <= imax; b += T) {
      double ai = w * b + bmin, bi = w * (b + 1) + bmin;
      sa += (double)obs[b] * (ai - mu); sb += (double)obs[b] * (bi - mu);
    }
    sa = block_reduce<double>(sa, shd, [](double a, double b) { return a + b; });
    sb = block_reduce<double>(sb, shd, [](double a, double b) { return a + b; });
    lambda = 1. / w * (log(sb) - log(sa));
  } else {
    double cnt = 0., mean = 0., lsum = 0.;
    for (int b = (int)max(cmin, (long long)imin) + tid; b <= imax; b += T) {
      if (!obs[b]) continue;
      double x = 0.5 * ((w * b + bmin) + (w * (b + 1) + bmin)) - mu;
      if (x <= 0.) continue;
      cnt += (double)obs[b]; mean += (double)obs[b] * x; lsum += (double)obs[b] * log(x);
    }
    cnt  = block_reduce<double>(cnt, shd, [](double a, double b) { return a + b; });
    mean = block_reduce<double>(mean, shd, [](double a, double b) { return a + b; });
    lsum = block_reduce<double>(lsum, shd, [](double a, double b) { return a + b; });
    if (cnt > 0.) {
      mean /= cnt; lsum /= cnt;
      double s = log(mean) - lsum;
      if (s > 1e-12) {
        double lo = 1e-6, hi = 1e6;
        for (int it = 0; it < 200; it++) {
          double mid = sqrt(lo * hi);
          double f = log(mid) - d_digamma(mid) - s;
          if (f > 0) lo = mid; else hi = mid;
          if (hi / lo < 1. + 1e-14) break;
        }
        tau = sqrt(lo * hi);
        lambda = tau / mean;
      } else { lambda = INFINITY; tau = INFINITY; }
    }
  }
  // survival fit: survfit[b] = newmass * S(UBound(b)) for b in [cmin, 2 nb)
  const bool have = !isinf(lambda);
  for (int b = tid; survfit && b < 2 * nb; b += T) {
    double v = 0.;
    if (have && b >= cmin) {
      double x = w * (b + 1) + bmin;
      double sv = (x < mu) ? 1.0 : (doexp ? exp(-lambda * (x - mu)) : d_gammq(tau, lambda * (x - mu)));
      v = newmass * sv;
    }
    survfit[b] = v;
  }
  if (tid == 0) {
    out[0] = mu; out[1] = lambda; out[2] = tau; out[3] = newmass; out[4] = pmass;
    out[5] = phi; out[6] = (double)cmin; out[7] = (double)kept;
  }
}

