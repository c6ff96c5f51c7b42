// rsg_mlp.hip -- the learned null-tail surrogate: a small MLP (GELU hidden layers, linear
// output) with hand-written forward, backward and Adam kernels.  No reference counterpart.
//
// One work group of 4 waves owns 32 samples.  Every layer product is a sequence of 32x32
// MFMA tiles: BF16 mode feeds v_mfma_f32_32x32x16_bf16 with BF16 weights and BF16-rounded
// activations and accumulates in FP32; FP32 mode feeds v_mfma_f32_32x32x2f32 (used by the
// gradient check).  Activations, pre-activations and back-propagated errors of the group's
// 32 samples stay in LDS for the whole forward + backward pass (activations are recomputed
// from the pre-activations where an operand needs them).  Per-group weight gradients
// go to a partials buffer that one kernel reduces in a fixed order (deterministic) before
// the Adam update, which also refreshes the BF16 weight copies.
//
// Parameter layout (FP32 master copy): for each layer l, W_l (dims[l+1] x dims[l], row
// major) then b_l (dims[l+1]).
#include "rsg_internal.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#define MLP_MAXL 6
#define MLP_MAXD 256
#define MLP_SB   32          // samples per work group

typedef float  v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

struct rsg_mlp {
  int      nl = 0;
  int      dims[MLP_MAXL + 1] = {};
  int64_t  offW[MLP_MAXL] = {}, offB[MLP_MAXL] = {};
  int64_t  P = 0;                          // parameter count
  int      bf16 = 1;
  float   *d_p = nullptr, *d_m = nullptr, *d_v = nullptr, *d_g = nullptr;
  uint16_t *d_wb = nullptr;                // BF16 copy of the whole parameter vector
  float   *d_part = nullptr;  int64_t part_blocks = 0;
  float   *d_io = nullptr;    size_t io_cap = 0;     // forward inputs + outputs, grown on demand
  float   *h_io = nullptr;                           // pinned host staging of d_io, same capacity
  int64_t  step = 0;                       // Adam step counter
  size_t   lds_bytes = 0;
  float   *d_scale = nullptr;                // input mean | input 1/std | output mean | output std
  std::vector<float> xmin, xmax;             // training range (empty: unbounded)
};

namespace {

struct MlpArgs {
  int nl, dims[MLP_MAXL + 1];
  int64_t offW[MLP_MAXL], offB[MLP_MAXL], P;
  int ldsZ[MLP_MAXL], ldsX, ldsG;             // float offsets in LDS
  const float *p; const uint16_t *wb;
  const float *xm, *xi, *ym, *ys;            // scalers: x' = (x - xm) * xi, y = y' * ys + ym
  const float *x, *t; const int *idx;
  int64_t nrows, first;                       // samples of this launch: rows first .. first+nrows-1 of idx/x
  float *y, *part;
  float scale;                                // dLoss/dy = scale * (y - t)
};

__host__ __device__ inline int ldl(int d) { return d | 1; }     // odd LDS row stride: no bank conflicts on columns

__device__ inline float gelu(float z)  { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ inline float dgelu(float z)
{
  return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * __expf(-0.5f * z * z) * 0.39894228040143268f;
}

__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// acc += A(32 x K) * B(K x 32) for one 32x32 tile.  fa(m, k), fb(k, n) return FP32 elements.
// Lane l supplies A row l%32 and B column l%32; the result element acc[r] of lane l is
// D[8*(r/4) + 4*(l/32) + r%4][l%32].
template <bool BF, class FA, class FB>
__device__ inline void mm_tile(v16f &acc, int K, FA fa, FB fb)
{
  const int l = threadIdx.x & 63, r32 = l & 31, hi = l >> 5;
  if constexpr (BF) {
    for (int k0 = 0; k0 < K; k0 += 16) {
      v8bf a, b;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int k = k0 + 8 * hi + e;
        a[e] = (__bf16)(k < K ? fa(r32, k) : 0.f);
        b[e] = (__bf16)(k < K ? fb(k, r32) : 0.f);
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
  } else {
    for (int k0 = 0; k0 < K; k0 += 2) {
      const int k = k0 + hi;
      const float a = k < K ? fa(r32, k) : 0.f, b = k < K ? fb(k, r32) : 0.f;
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
    }
  }
}

template <bool BF>
__device__ inline float weight(const MlpArgs &a, int l, int o, int i)
{
  const int64_t q = a.offW[l] + (int64_t)o * a.dims[l] + i;
  if constexpr (BF) return bf2f(a.wb[q]);
  else return a.p[q];
}

// Forward (and, with BWD, backward) pass of MLP_SB samples per work group.
template <bool BF, bool BWD>
__global__ void __launch_bounds__(256) k_mlp(MlpArgs a)
{
  extern __shared__ float sm[];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, r32 = lane & 31, hi = lane >> 5;
  const int64_t s0 = (int64_t)blockIdx.x * MLP_SB;
  const int nvalid = (int)min((int64_t)MLP_SB, a.nrows - s0);
  const int nl = a.nl, d0 = a.dims[0], dl = a.dims[nl];
  float *X = sm + a.ldsX, *G = sm + a.ldsG;

  // inputs of the group's samples (rows past the end are zero)
  for (int e = tid; e < MLP_SB * d0; e += 256) {
    const int s = e / d0, i = e % d0;
    float v = 0.f;
    if (s < nvalid) {
      const int64_t row = a.idx ? a.idx[a.first + s0 + s] : a.first + s0 + s;
      v = (a.x[row * d0 + i] - a.xm[i]) * a.xi[i];
    }
    X[s * ldl(d0) + i] = v;
  }
  __syncthreads();

  // ---- forward ----  (layer l reads X or GELU(Z_{l-1}), writes Z_l, or G for the output)
  for (int l = 0; l < nl; l++) {
    const int din = a.dims[l], dout = a.dims[l + 1];
    const float *Zin = (l == 0) ? X : sm + a.ldsZ[l - 1];
    const int ldi = ldl(din);
    float *Zo = (l < nl - 1) ? sm + a.ldsZ[l] : G;
    const int ldo = ldl(dout);
    for (int tn = wave; tn * 32 < dout; tn += 4) {
      v16f acc = {};
      const int o = tn * 32 + r32;
      mm_tile<BF>(acc, din,
                  [&](int m, int k) { const float z = Zin[m * ldi + k]; return l == 0 ? z : gelu(z); },
                  [&](int k, int n) { return o < dout ? weight<BF>(a, l, o, k) : 0.f; });
      if (o < dout) {
        const float bias = a.p[a.offB[l] + o];
#pragma unroll
        for (int r = 0; r < 16; r++) {
          const int s = 8 * (r >> 2) + 4 * hi + (r & 3);
          Zo[s * ldo + o] = acc[r] + bias;
        }
      }
    }
    __syncthreads();
  }

  if constexpr (!BWD) {
    for (int e = tid; e < nvalid * dl; e += 256) {
      const int s = e / dl, o = e % dl;
      a.y[(s0 + s) * dl + o] = G[s * ldl(dl) + o] * a.ys[o] + a.ym[o];
    }
    return;
  } else {
    float *part = a.part + (int64_t)blockIdx.x * (a.P + 1);
    // output error and the group's loss partial
    __shared__ float red[4];
    float ls = 0.f;
    for (int e = tid; e < MLP_SB * dl; e += 256) {
      const int s = e / dl, o = e % dl;
      float g = 0.f;
      if (s < nvalid) {
        const int64_t row = a.idx ? a.idx[a.first + s0 + s] : a.first + s0 + s;
        const float d = G[s * ldl(dl) + o] - (a.t[row * dl + o] - a.ym[o]) / a.ys[o];
        ls += d * d;
        g = a.scale * d;
      }
      G[s * ldl(dl) + o] = g;
    }
    for (int off = 32; off > 0; off >>= 1) ls += __shfl_down(ls, off, 64);
    if (lane == 0) red[wave] = ls;
    __syncthreads();
    if (tid == 0) part[a.P] = red[0] + red[1] + red[2] + red[3];

    // ---- backward ----
    for (int l = nl - 1; l >= 0; l--) {
      const int din = a.dims[l], dout = a.dims[l + 1];
      const int ldo = ldl(dout), ldi = ldl(din);
      const float *Gl = (l == nl - 1) ? G : sm + a.ldsZ[l];
      float *Zin = (l == 0) ? X : sm + a.ldsZ[l - 1];
      // dW_l[o][i] = sum_s G[s][o] Hin[s][i], Hin = X or GELU(Z_{l-1})
      const int nto = (dout + 31) / 32, nti = (din + 31) / 32;
      for (int t = wave; t < nto * nti; t += 4) {
        const int to = t / nti, ti = t % nti;
        v16f acc = {};
        mm_tile<BF>(acc, MLP_SB,
                    [&](int m, int k) { const int o = to * 32 + m; return o < dout ? Gl[k * ldo + o] : 0.f; },
                    [&](int k, int n) {
                      const int i = ti * 32 + n;
                      if (i >= din) return 0.f;
                      const float z = Zin[k * ldi + i];
                      return l == 0 ? z : gelu(z);
                    });
        const int i = ti * 32 + r32;
        if (i < din) {
#pragma unroll
          for (int r = 0; r < 16; r++) {
            const int o = to * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
            if (o < dout) part[a.offW[l] + (int64_t)o * din + i] = acc[r];
          }
        }
      }
      for (int o = tid; o < dout; o += 256) {
        float sacc = 0.f;
        for (int s = 0; s < MLP_SB; s++) sacc += Gl[s * ldo + o];
        part[a.offB[l] + o] = sacc;
      }
      if (l > 0) {
        __syncthreads();                 // every wave is done reading Z_{l-1} for dW
        // G_{l-1}[s][i] = (sum_o G[s][o] W[o][i]) * GELU'(Z_{l-1}[s][i]), written over Z_{l-1}
        for (int tn = wave; tn * 32 < din; tn += 4) {
          v16f acc = {};
          const int i = tn * 32 + r32;
          mm_tile<BF>(acc, dout,
                      [&](int m, int k) { return Gl[m * ldo + k]; },
                      [&](int k, int n) { return i < din ? weight<BF>(a, l, k, i) : 0.f; });
          if (i < din) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
              const int s = 8 * (r >> 2) + 4 * hi + (r & 3);
              Zin[s * ldi + i] = acc[r] * dgelu(Zin[s * ldi + i]);
            }
          }
        }
      }
      __syncthreads();
    }
  }
}

// g[p] = sum over groups of the partials (fixed order); then, with ADAM, the Adam update
// and the BF16 refresh.  The loss sum lands in loss[0] when loss is given.
template <bool ADAM>
__global__ void k_mlp_reduce(const float *__restrict__ part, int64_t nblk, int64_t P, float *__restrict__ g,
                             float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                             uint16_t *__restrict__ wb, float lr_t, float b1, float b2, float eps,
                             double *__restrict__ loss)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q > P) return;
  float acc = 0.f;
  for (int64_t b = 0; b < nblk; b++) acc += part[b * (P + 1) + q];
  if (q == P) { if (loss) loss[0] = acc; return; }
  if (g) g[q] = acc;
  if constexpr (ADAM) {
    const float mm = b1 * m[q] + (1.f - b1) * acc;
    const float vv = b2 * v[q] + (1.f - b2) * acc * acc;
    m[q] = mm; v[q] = vv;
    const float np = p[q] - lr_t * mm / (sqrtf(vv) + eps);
    p[q] = np;
    wb[q] = (uint16_t)(__float_as_uint((float)(__bf16)np) >> 16);
  }
}

__global__ void k_bf16_copy(const float *__restrict__ p, int64_t P, uint16_t *__restrict__ wb)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < P) wb[q] = (uint16_t)(__float_as_uint((float)(__bf16)p[q]) >> 16);
}

MlpArgs make_args(const rsg_mlp *m)
{
  MlpArgs a;
  memset(&a, 0, sizeof(a));
  a.nl = m->nl;
  for (int l = 0; l <= m->nl; l++) a.dims[l] = m->dims[l];
  for (int l = 0; l < m->nl; l++) { a.offW[l] = m->offW[l]; a.offB[l] = m->offB[l]; }
  a.P = m->P;
  int off = 0;
  a.ldsX = off; off += MLP_SB * ldl(m->dims[0]);
  for (int l = 0; l + 1 < m->nl; l++) { a.ldsZ[l] = off; off += MLP_SB * ldl(m->dims[l + 1]); }
  a.ldsG = off; off += MLP_SB * ldl(m->dims[m->nl]);
  a.p = m->d_p; a.wb = m->d_wb;
  const int d0 = m->dims[0], dl = m->dims[m->nl];
  a.xm = m->d_scale; a.xi = m->d_scale + d0; a.ym = m->d_scale + 2 * d0; a.ys = m->d_scale + 2 * d0 + dl;
  return a;
}

size_t lds_bytes(const rsg_mlp *m)
{
  MlpArgs a = make_args(m);
  return sizeof(float) * (size_t)(a.ldsG + MLP_SB * ldl(m->dims[m->nl]));
}

int launch_mlp(rsg_ctx *c, rsg_mlp *m, const MlpArgs &a, bool bwd, char *errbuf)
{
  const unsigned nblk = (unsigned)((a.nrows + MLP_SB - 1) / MLP_SB);
  if (nblk == 0) return RSG_OK;
  const size_t lds = m->lds_bytes;
  if (m->bf16) {
    if (bwd) hipLaunchKernelGGL((k_mlp<true, true>), dim3(nblk), dim3(256), lds, c->stream, a);
    else     hipLaunchKernelGGL((k_mlp<true, false>), dim3(nblk), dim3(256), lds, c->stream, a);
  } else {
    if (bwd) hipLaunchKernelGGL((k_mlp<false, true>), dim3(nblk), dim3(256), lds, c->stream, a);
    else     hipLaunchKernelGGL((k_mlp<false, false>), dim3(nblk), dim3(256), lds, c->stream, a);
  }
  HIPCHK(hipGetLastError());
  return RSG_OK;
}

int ensure_part(rsg_mlp *m, int64_t nblk, char *errbuf)
{
  if (m->part_blocks >= nblk) return RSG_OK;
  if (m->d_part) hipFree(m->d_part);
  m->d_part = nullptr; m->part_blocks = 0;
  HIPCHK(hipMalloc(&m->d_part, sizeof(float) * (size_t)(nblk * (m->P + 1))));
  m->part_blocks = nblk;
  return RSG_OK;
}

} // namespace

extern "C" int rsg_mlp_create(rsg_ctx *c, int nl, const int *dims, uint64_t seed, rsg_mlp **ret, char *errbuf)
{
  *ret = nullptr;
  if (!c || !dims || nl < 1 || nl > MLP_MAXL) return rsg_fail(errbuf, "rsg_mlp_create: 1..%d layers", MLP_MAXL);
  for (int l = 0; l <= nl; l++)
    if (dims[l] < 1 || dims[l] > MLP_MAXD) return rsg_fail(errbuf, "rsg_mlp_create: layer widths must be 1..%d", MLP_MAXD);
  HIPCHK(hipSetDevice(c->device));
  rsg_mlp *m = new rsg_mlp();
  m->nl = nl;
  for (int l = 0; l <= nl; l++) m->dims[l] = dims[l];
  int64_t off = 0;
  for (int l = 0; l < nl; l++) {
    m->offW[l] = off; off += (int64_t)dims[l] * dims[l + 1];
    m->offB[l] = off; off += dims[l + 1];
  }
  m->P = off;
  m->lds_bytes = lds_bytes(m);
  int maxlds = 0;
  hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
  if ((size_t)maxlds < m->lds_bytes) {
    delete m;
    return rsg_fail(errbuf, "rsg_mlp_create: needs %zu bytes of LDS per work group, device allows %d", lds_bytes(m), maxlds);
  }
  // Glorot-uniform weights from Philox (seed, stream 0x4D4C50, layer, index), zero biases
  std::vector<float> h((size_t)m->P, 0.f);
  for (int l = 0; l < nl; l++) {
    const double lim = sqrt(6.0 / (double)(dims[l] + dims[l + 1]));
    for (int64_t q = 0; q < (int64_t)dims[l] * dims[l + 1]; q++) {
      const double u = (rsg_philox(seed, 0x4D4C50u, (uint32_t)l, (uint32_t)q, 0) + 0.5) * (1.0 / 4294967296.0);
      h[(size_t)(m->offW[l] + q)] = (float)((2.0 * u - 1.0) * lim);
    }
  }
  const size_t pb = sizeof(float) * (size_t)m->P;
  if (hipMalloc(&m->d_p, pb) != hipSuccess || hipMalloc(&m->d_m, pb) != hipSuccess ||
      hipMalloc(&m->d_v, pb) != hipSuccess || hipMalloc(&m->d_g, pb) != hipSuccess ||
      hipMalloc(&m->d_wb, sizeof(uint16_t) * (size_t)m->P) != hipSuccess ||
      hipMalloc(&m->d_scale, sizeof(float) * 2 * (size_t)(dims[0] + dims[nl])) != hipSuccess ||
      hipMemcpy(m->d_p, h.data(), pb, hipMemcpyHostToDevice) != hipSuccess) {
    rsg_mlp_destroy(c, m);
    return rsg_fail(errbuf, "rsg_mlp_create: device allocation failed");
  }
  if (rsg_mlp_set_params(c, m, nullptr, nullptr, errbuf) != RSG_OK ||
      rsg_mlp_set_scaling(c, m, nullptr, nullptr, nullptr, nullptr, errbuf) != RSG_OK) { rsg_mlp_destroy(c, m); return RSG_FAIL; }
  *ret = m;
  return RSG_OK;
}

extern "C" void rsg_mlp_destroy(rsg_ctx *c, rsg_mlp *m)
{
  if (!m) return;
  if (c) { hipSetDevice(c->device); hipStreamSynchronize(c->stream); }
  hipFree(m->d_p); hipFree(m->d_m); hipFree(m->d_v); hipFree(m->d_g); hipFree(m->d_wb); hipFree(m->d_scale);
  if (m->d_part) hipFree(m->d_part);
  if (m->d_io) hipFree(m->d_io);
  if (m->h_io) hipHostFree(m->h_io);
  delete m;
}

extern "C" int rsg_mlp_set_scaling(rsg_ctx *c, rsg_mlp *m, const float *xmean, const float *xstd,
                                   const float *ymean, const float *ystd, char *errbuf)
{
  if (!c || !m) return rsg_fail(errbuf, "rsg_mlp_set_scaling: bad arguments");
  const int d0 = m->dims[0], dl = m->dims[m->nl];
  std::vector<float> h((size_t)(2 * (d0 + dl)));
  for (int i = 0; i < d0; i++) {
    h[(size_t)i] = xmean ? xmean[i] : 0.f;
    const float sd = xstd ? xstd[i] : 1.f;
    if (!(sd > 0.f)) return rsg_fail(errbuf, "rsg_mlp_set_scaling: input std %d must be > 0", i);
    h[(size_t)(d0 + i)] = 1.f / sd;
  }
  for (int o = 0; o < dl; o++) {
    h[(size_t)(2 * d0 + o)] = ymean ? ymean[o] : 0.f;
    const float sd = ystd ? ystd[o] : 1.f;
    if (!(sd > 0.f)) return rsg_fail(errbuf, "rsg_mlp_set_scaling: output std %d must be > 0", o);
    h[(size_t)(2 * d0 + dl + o)] = sd;
  }
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpy(m->d_scale, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
  return RSG_OK;
}

extern "C" int rsg_mlp_set_range(rsg_mlp *m, const float *xmin, const float *xmax)
{
  if (!m) return RSG_FAIL;
  if (!xmin || !xmax) { m->xmin.clear(); m->xmax.clear(); return RSG_OK; }
  m->xmin.assign(xmin, xmin + m->dims[0]);
  m->xmax.assign(xmax, xmax + m->dims[0]);
  return RSG_OK;
}

extern "C" int rsg_mlp_in_range(const rsg_mlp *m, const float *x)
{
  if (!m || !x) return 0;
  for (size_t i = 0; i < m->xmin.size(); i++)
    if (!(x[i] >= m->xmin[i] && x[i] <= m->xmax[i])) return 0;
  return 1;
}

extern "C" int rsg_mlp_set_precision(rsg_mlp *m, int bf16)
{
  if (!m) return RSG_FAIL;
  m->bf16 = bf16 ? 1 : 0;
  return RSG_OK;
}

// Uploads W/b (W == NULL keeps the device values), resets the Adam state and refreshes the
// BF16 weight copy.
extern "C" int rsg_mlp_set_params(rsg_ctx *c, rsg_mlp *m, const float *const *W, const float *const *b, char *errbuf)
{
  if (!c || !m) return rsg_fail(errbuf, "rsg_mlp_set_params: bad arguments");
  HIPCHK(hipSetDevice(c->device));
  const size_t pb = sizeof(float) * (size_t)m->P;
  if (W) {
    std::vector<float> h((size_t)m->P, 0.f);
    for (int l = 0; l < m->nl; l++) {
      memcpy(&h[(size_t)m->offW[l]], W[l], sizeof(float) * (size_t)m->dims[l] * m->dims[l + 1]);
      if (b) memcpy(&h[(size_t)m->offB[l]], b[l], sizeof(float) * (size_t)m->dims[l + 1]);
    }
    HIPCHK(hipMemcpyAsync(m->d_p, h.data(), pb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  HIPCHK(hipMemsetAsync(m->d_m, 0, pb, c->stream));
  HIPCHK(hipMemsetAsync(m->d_v, 0, pb, c->stream));
  m->step = 0;
  hipLaunchKernelGGL(k_bf16_copy, dim3((unsigned)((m->P + 255) / 256)), dim3(256), 0, c->stream, m->d_p, m->P, m->d_wb);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return RSG_OK;
}

extern "C" int rsg_mlp_get_params(rsg_ctx *c, rsg_mlp *m, float *const *W, float *const *b, char *errbuf)
{
  if (!c || !m || !W) return rsg_fail(errbuf, "rsg_mlp_get_params: bad arguments");
  HIPCHK(hipSetDevice(c->device));
  std::vector<float> h((size_t)m->P);
  HIPCHK(hipMemcpyAsync(h.data(), m->d_p, sizeof(float) * (size_t)m->P, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int l = 0; l < m->nl; l++) {
    memcpy(W[l], &h[(size_t)m->offW[l]], sizeof(float) * (size_t)m->dims[l] * m->dims[l + 1]);
    if (b) memcpy(b[l], &h[(size_t)m->offB[l]], sizeof(float) * (size_t)m->dims[l + 1]);
  }
  return RSG_OK;
}

extern "C" int rsg_mlp_forward(rsg_ctx *c, rsg_mlp *m, const float *x, int64_t B, float *y, char *errbuf)
{
  if (!c || !m || !x || !y || B < 0) return rsg_fail(errbuf, "rsg_mlp_forward: bad arguments");
  if (B == 0) return RSG_OK;
  HIPCHK(hipSetDevice(c->device));
  const size_t xb = sizeof(float) * (size_t)(B * m->dims[0]), yb = sizeof(float) * (size_t)(B * m->dims[m->nl]);
  if (m->io_cap < xb + yb) {
    if (m->d_io) hipFree(m->d_io);
    if (m->h_io) hipHostFree(m->h_io);
    m->d_io = m->h_io = nullptr; m->io_cap = 0;
    HIPCHK(hipMalloc(&m->d_io, xb + yb));
    HIPCHK(hipHostMalloc((void **)&m->h_io, xb + yb, hipHostMallocDefault));
    m->io_cap = xb + yb;
  }
  // inputs and outputs go through pinned staging: small pageable copies cost tens of us each
  float *d_x = m->d_io, *d_y = m->d_io + (size_t)(B * m->dims[0]);
  float *h_x = m->h_io, *h_y = m->h_io + (size_t)(B * m->dims[0]);
  memcpy(h_x, x, xb);
  MlpArgs a = make_args(m);
  a.x = d_x; a.y = d_y; a.nrows = B; a.first = 0;
  int st = RSG_OK;
  if (hipMemcpyAsync(d_x, h_x, xb, hipMemcpyHostToDevice, c->stream) != hipSuccess) st = rsg_fail(errbuf, "rsg_mlp_forward: upload failed");
  int t; rsg_tstart(c, "mlp_forward", &t);
  if (st == RSG_OK) st = launch_mlp(c, m, a, false, errbuf);
  rsg_tstop(c, t);
  if (st == RSG_OK && (hipMemcpyAsync(h_y, d_y, yb, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                       hipStreamSynchronize(c->stream) != hipSuccess))
    st = rsg_fail(errbuf, "rsg_mlp_forward: download failed");
  if (st == RSG_OK) memcpy(y, h_y, yb);
  return st;
}

extern "C" int rsg_mlp_grad(rsg_ctx *c, rsg_mlp *m, const float *x, const float *t, int64_t B,
                            float *const *dW, float *const *db, double *loss, char *errbuf)
{
  if (!c || !m || !x || !t || B < 1 || !dW || !db) return rsg_fail(errbuf, "rsg_mlp_grad: bad arguments");
  HIPCHK(hipSetDevice(c->device));
  const int64_t nblk = (B + MLP_SB - 1) / MLP_SB;
  if (ensure_part(m, nblk, errbuf) != RSG_OK) return RSG_FAIL;
  float *d_x = nullptr, *d_t = nullptr; double *d_l = nullptr;
  const size_t xb = sizeof(float) * (size_t)(B * m->dims[0]), tb = sizeof(float) * (size_t)(B * m->dims[m->nl]);
  HIPCHK(hipMalloc(&d_x, xb));
  if (hipMalloc(&d_t, tb) != hipSuccess || hipMalloc(&d_l, sizeof(double)) != hipSuccess) {
    hipFree(d_x); hipFree(d_t);
    return rsg_fail(errbuf, "rsg_mlp_grad: out of device memory");
  }
  MlpArgs a = make_args(m);
  a.x = d_x; a.t = d_t; a.nrows = B; a.first = 0; a.part = m->d_part;
  a.scale = (float)(2.0 / ((double)B * m->dims[m->nl]));
  int st = RSG_OK;
  if (hipMemcpyAsync(d_x, x, xb, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(d_t, t, tb, hipMemcpyHostToDevice, c->stream) != hipSuccess) st = rsg_fail(errbuf, "rsg_mlp_grad: upload failed");
  if (st == RSG_OK) st = launch_mlp(c, m, a, true, errbuf);
  if (st == RSG_OK) {
    hipLaunchKernelGGL((k_mlp_reduce<false>), dim3((unsigned)((m->P + 1 + 255) / 256)), dim3(256), 0, c->stream,
                       m->d_part, nblk, m->P, m->d_g, nullptr, nullptr, nullptr, nullptr, 0.f, 0.f, 0.f, 0.f, d_l);
    if (hipGetLastError() != hipSuccess) st = rsg_fail(errbuf, "rsg_mlp_grad: reduce launch failed");
  }
  std::vector<float> g((size_t)m->P);
  double lsd = 0.;
  if (st == RSG_OK && (hipMemcpyAsync(g.data(), m->d_g, sizeof(float) * g.size(), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                       hipMemcpyAsync(&lsd, d_l, sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                       hipStreamSynchronize(c->stream) != hipSuccess))
    st = rsg_fail(errbuf, "rsg_mlp_grad: download failed");
  if (st == RSG_OK) {
    for (int l = 0; l < m->nl; l++) {
      memcpy(dW[l], &g[(size_t)m->offW[l]], sizeof(float) * (size_t)m->dims[l] * m->dims[l + 1]);
      memcpy(db[l], &g[(size_t)m->offB[l]], sizeof(float) * (size_t)m->dims[l + 1]);
    }
    if (loss) *loss = lsd / ((double)B * m->dims[m->nl]);
  }
  hipFree(d_x); hipFree(d_t); hipFree(d_l);
  return st;
}

extern "C" int rsg_mlp_train(rsg_ctx *c, rsg_mlp *m, const float *x, const float *t, int64_t n,
                             int epochs, int batch, double lr, uint64_t seed, double *final_loss, char *errbuf)
{
  if (!c || !m || !x || !t || n < 1 || epochs < 1 || batch < 1) return rsg_fail(errbuf, "rsg_mlp_train: bad arguments");
  HIPCHK(hipSetDevice(c->device));
  if (batch > n) batch = (int)n;
  const int64_t steps = (n + batch - 1) / batch;
  const int64_t nblk = (batch + MLP_SB - 1) / MLP_SB;
  if (ensure_part(m, nblk, errbuf) != RSG_OK) return RSG_FAIL;
  const size_t xb = sizeof(float) * (size_t)(n * m->dims[0]), tb = sizeof(float) * (size_t)(n * m->dims[m->nl]);
  float *d_x = nullptr, *d_t = nullptr; int *d_idx = nullptr; double *d_loss = nullptr;
  if (hipMalloc(&d_x, xb) != hipSuccess || hipMalloc(&d_t, tb) != hipSuccess ||
      hipMalloc(&d_idx, sizeof(int) * (size_t)n) != hipSuccess ||
      hipMalloc(&d_loss, sizeof(double) * (size_t)steps) != hipSuccess) {
    hipFree(d_x); hipFree(d_t); hipFree(d_idx); hipFree(d_loss);
    return rsg_fail(errbuf, "rsg_mlp_train: out of device memory");
  }
  int st = RSG_OK;
  if (hipMemcpyAsync(d_x, x, xb, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(d_t, t, tb, hipMemcpyHostToDevice, c->stream) != hipSuccess) st = rsg_fail(errbuf, "rsg_mlp_train: upload failed");
  std::vector<int> perm((size_t)n);
  std::vector<double> losses((size_t)steps);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  int tm; rsg_tstart(c, "mlp_train", &tm);
  for (int ep = 0; ep < epochs && st == RSG_OK; ep++) {
    // Fisher-Yates with Philox draws keyed by (seed, epoch, position)
    for (int64_t k = 0; k < n; k++) perm[(size_t)k] = (int)k;
    for (int64_t k = n - 1; k > 0; k--) {
      const uint64_t r = ((uint64_t)rsg_philox(seed, 0x54524Eu, (uint32_t)ep, (uint32_t)k, 0) << 32) |
                         rsg_philox(seed, 0x54524Eu, (uint32_t)ep, (uint32_t)k, 1);
      const int64_t j = (int64_t)(r % (uint64_t)(k + 1));
      std::swap(perm[(size_t)k], perm[(size_t)j]);
    }
    if (hipMemcpyAsync(d_idx, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) { st = rsg_fail(errbuf, "rsg_mlp_train: index upload failed"); break; }
    for (int64_t sidx = 0; sidx < steps && st == RSG_OK; sidx++) {
      const int64_t first = sidx * batch, rows = std::min<int64_t>(batch, n - first);
      MlpArgs a = make_args(m);
      a.x = d_x; a.t = d_t; a.idx = d_idx; a.first = first; a.nrows = rows; a.part = m->d_part;
      a.scale = (float)(2.0 / ((double)rows * m->dims[m->nl]));
      st = launch_mlp(c, m, a, true, errbuf);
      if (st != RSG_OK) break;
      m->step++;
      const double bc1 = 1.0 - pow((double)b1, (double)m->step), bc2 = 1.0 - pow((double)b2, (double)m->step);
      const float lr_t = (float)(lr * sqrt(bc2) / bc1);
      hipLaunchKernelGGL((k_mlp_reduce<true>), dim3((unsigned)((m->P + 1 + 255) / 256)), dim3(256), 0, c->stream,
                         m->d_part, (rows + MLP_SB - 1) / MLP_SB, m->P, (float *)nullptr, m->d_p, m->d_m, m->d_v, m->d_wb,
                         lr_t, b1, b2, eps, d_loss + sidx);
      if (hipGetLastError() != hipSuccess) st = rsg_fail(errbuf, "rsg_mlp_train: update launch failed");
    }
  }
  rsg_tstop(c, tm);
  if (st == RSG_OK && (hipMemcpyAsync(losses.data(), d_loss, sizeof(double) * (size_t)steps, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                       hipStreamSynchronize(c->stream) != hipSuccess))
    st = rsg_fail(errbuf, "rsg_mlp_train: download failed");
  if (st == RSG_OK && final_loss) {
    double s = 0.;
    for (int64_t k = 0; k < steps; k++) s += losses[(size_t)k];
    *final_loss = s / ((double)n * m->dims[m->nl]);
  }
  hipFree(d_x); hipFree(d_t); hipFree(d_idx); hipFree(d_loss);
  return st;
}
