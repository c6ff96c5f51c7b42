// rsg_gram.hip -- pair-count tables as one-hot Gram contractions on CDNA4 MFMA.
//
// Replaces the O(L^2 N) gather loop of mutual_naive_ppij (src/correlators.c:1695-1780,
// driven by corr_NaivePP :1300-1317).  With GAPASCHAR == 0 (src/rscape_config.h:48) the
// table of a column pair is 4 x 4: c[x][y] = sum_s w_s [a_si == x][a_sj == y] over
// canonical symbols.  Writing X for the N x 4L one-hot matrix, all tables at once are
// the Gram C = X^T diag(w) X; only the upper-triangular column blocks are formed.
//
// Unit weights: int8 MFMA (v_mfma_i32_32x32x32_i8), int32 accumulation, bit-exact.
//   The alignment is staged as one byte per (column, sequence) holding 1 << symbol
//   (0 for gap / N).  A lane that feeds feature (column, symbol s) masks the byte with
//   0x01 << s, so the operand byte is 2^s or 0; the product of the two operands of a
//   pair cell (x, y) is 2^(x+y), which the epilogue shifts back out.  One staged byte
//   therefore feeds all four symbol rows, and LDS / L2 traffic is 1 byte per
//   (column, sequence) instead of 4.
// Weights: split-BF16 (w = hi + lo [+ lo2]) MFMA (v_mfma_f32_32x32x16_bf16), FP32
//   accumulation of all split terms into one accumulator.
//
// Output: one 16-entry record per column pair (i < j) at cnt[(i*L + j)*16 + x*4 + y],
// the layout of pp[i][j][IDX(x,y,4)] in the reference.
#include "rsg_internal.h"
#include <map>
#include <tuple>
#include <algorithm>

typedef int   v4i  __attribute__((ext_vector_type(4)));
typedef int   v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

#define KT 64            // sequences per LDS stage of the int8 Gram

int64_t rsg_npad(int64_t N) { return ((N + KT - 1) / KT) * KT; }

// ------------------------------------------------------------------------------------
// Pack: ax (N x L, row-major) -> axT (Lpad x Npad raw codes) and xt (Lpad x Npad nibbles).
// 64 x 64 tiles through LDS so both the read and the write are coalesced.
__global__ void k_pack(const uint8_t *__restrict__ ax, int64_t N, int64_t L, int64_t Npad,
                       uint8_t *__restrict__ axT, uint8_t *__restrict__ xt)
{
  __shared__ uint8_t tile[64][65];
  int64_t s0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
  int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;      // 256 threads: 64 x 4
  for (int r = ty; r < 64; r += 4) {
    int64_t s = s0 + r, c = c0 + tx;
    tile[r][tx] = (s < N && c < L) ? ax[s * L + c] : (uint8_t)4;
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    int64_t c = c0 + r, s = s0 + tx;
    uint8_t a = tile[tx][r];
    axT[c * Npad + s] = a;
    xt[c * Npad + s]  = (a < 4) ? (uint8_t)(1u << a) : (uint8_t)0;
  }
}

int rsg_pack(rsg_ctx *c, rsg_msa *m, char *errbuf)
{
  dim3 grid((unsigned)(m->Lpad / 64), (unsigned)(m->Npad / 64));
  int t; rsg_tstart(c, "pack", &t);
  hipLaunchKernelGGL(k_pack, grid, dim3(256), 0, c->stream, m->d_ax, m->N, m->L, m->Npad, m->d_axT, m->d_xt);
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  return RSG_OK;
}

// ------------------------------------------------------------------------------------
// int8 Gram.  Work-group tile: BI columns (i side) x BJ columns (j side); WI x WJ waves,
// each wave TI x TJ MFMA tiles of 8 x 8 columns (32 x 32 features).
//
// MFMA 32x32 operand / result lane layout (CDNA3/4): A row m and B column n are held
// by lanes with lane % 32 == m (resp. n), the 32 K values split in two halves by
// lane / 32; result register r of lane l holds D[8*(r/4) + 4*(l/32) + r%4][l % 32].
// A rows are j-side features (m = 4*jj + y), B columns i-side features (n = 4*ii + x),
// so lane l owns (ii = (l%32)/4, x = l%4) and registers 4q..4q+3 hold y = 0..3 of
// jj = 2q + l/32.  A Gram uses the same K mapping for A and B, so the assignment of K
// values to lane halves need not be known.
template <int TI, int TJ, int WI, int WJ, int KIND>
__global__ void __launch_bounds__(64 * WI * WJ)
k_gram_i8(const uint8_t *__restrict__ xt, int64_t Npad, int64_t L, void *__restrict__ out,
          const int2 *__restrict__ tiles, int ntiles, int per_xcd)
{
  constexpr int BI = 8 * TI * WI, BJ = 8 * TJ * WJ, NT = 64 * WI * WJ;
  constexpr int CSA = BJ * 16 + 128, CSB = BI * 16 + 128;     // LDS chunk-row strides
  constexpr int STAGE = 4 * (CSA + CSB);
  constexpr int NCH = (BI + BJ) * (KT / 16);                   // 16-byte chunks per stage
  constexpr int NR = (NCH + NT - 1) / NT;
  __shared__ __attribute__((aligned(16))) uint8_t lds[2 * STAGE];

  // XCD-aware mapping: consecutive tiles of the (bi, bj)-sorted list land on one XCD.
  const int g = blockIdx.x;
  const int t = (g & 7) * per_xcd + (g >> 3);
  if (t >= ntiles) return;
  const int2 tl = tiles[t];
  const int64_t i0 = (int64_t)tl.x * BI, j0 = (int64_t)tl.y * BJ;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = wave % WI, wj = wave / WI;
  const unsigned mask = 0x01010101u << (lane & 3);
  const int fcol = (lane & 31) >> 2, khalf = lane >> 5;

  v16i acc[TI][TJ];
#pragma unroll
  for (int a = 0; a < TI; a++)
#pragma unroll
    for (int b = 0; b < TJ; b++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[a][b][r] = 0;

  uint4 pre[NR];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int r = 0; r < NR; r++) {
      int ch = tid + r * NT;
      if (ch < NCH) {
        int cc = ch >> 2, kc = ch & 3;
        int64_t col = (cc < BJ) ? (j0 + cc) : (i0 + cc - BJ);
        pre[r] = *(const uint4 *)(xt + col * Npad + k0 + kc * 16);
      }
    }
  };
  auto sstore = [&](int buf) {
    uint8_t *base = lds + buf * STAGE;
#pragma unroll
    for (int r = 0; r < NR; r++) {
      int ch = tid + r * NT;
      if (ch < NCH) {
        int cc = ch >> 2, kc = ch & 3;
        uint8_t *p = (cc < BJ) ? (base + kc * CSA + cc * 16) : (base + 4 * CSA + kc * CSB + (cc - BJ) * 16);
        *(uint4 *)p = pre[r];
      }
    }
  };

  const int64_t nkt = Npad / KT;
  gload(0);
  sstore(0);
  __syncthreads();
  for (int64_t kt = 0; kt < nkt; kt++) {
    const int buf = (int)(kt & 1);
    if (kt + 1 < nkt) gload((kt + 1) * KT);
    const uint8_t *A = lds + buf * STAGE;
    const uint8_t *B = A + 4 * CSA;
#pragma unroll
    for (int ks = 0; ks < KT / 32; ks++) {
      const int kc = ks * 2 + khalf;
      v4i af[TJ], bf[TI];
#pragma unroll
      for (int b = 0; b < TJ; b++) {
        uint4 v = *(const uint4 *)(A + kc * CSA + (wj * 8 * TJ + b * 8 + fcol) * 16);
        af[b] = (v4i){(int)(v.x & mask), (int)(v.y & mask), (int)(v.z & mask), (int)(v.w & mask)};
      }
#pragma unroll
      for (int a = 0; a < TI; a++) {
        uint4 v = *(const uint4 *)(B + kc * CSB + (wi * 8 * TI + a * 8 + fcol) * 16);
        bf[a] = (v4i){(int)(v.x & mask), (int)(v.y & mask), (int)(v.z & mask), (int)(v.w & mask)};
      }
#pragma unroll
      for (int a = 0; a < TI; a++)
#pragma unroll
        for (int b = 0; b < TJ; b++)
          acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[b], bf[a], acc[a][b], 0, 0, 0);
    }
    if (kt + 1 < nkt) sstore(buf ^ 1);
    __syncthreads();
  }

  // epilogue: lane owns (ii, x); undo the 2^(x+y) operand scaling
  const int x = lane & 3, ii = (lane & 31) >> 2;
#pragma unroll
  for (int a = 0; a < TI; a++) {
    const int64_t i = i0 + wi * 8 * TI + a * 8 + ii;
#pragma unroll
    for (int b = 0; b < TJ; b++) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int64_t j = j0 + wj * 8 * TJ + b * 8 + 2 * q + khalf;
        if (i < j && j < L) {
          const int64_t off = (i * L + j) * 16 + x * 4;
          unsigned c0 = (unsigned)acc[a][b][4 * q + 0] >> (x + 0);
          unsigned c1 = (unsigned)acc[a][b][4 * q + 1] >> (x + 1);
          unsigned c2 = (unsigned)acc[a][b][4 * q + 2] >> (x + 2);
          unsigned c3 = (unsigned)acc[a][b][4 * q + 3] >> (x + 3);
          if (KIND == CNT_U16) {
            *(uint2 *)((uint16_t *)out + off) = make_uint2(c0 | (c1 << 16), c2 | (c3 << 16));
          } else {
            *(uint4 *)((uint32_t *)out + off) = make_uint4(c0, c1, c2, c3);
          }
        }
      }
    }
  }
}

// Tile-shape table (the autotuner sweeps these).
struct GramShape { int TI, TJ, WI, WJ; const char *name; };
static const GramShape kShapes[] = {
  {4, 4, 2, 2, "64x64c/4w(4x4)"},
  {2, 4, 2, 2, "32x64c/4w(2x4)"},
  {4, 2, 2, 2, "64x32c/4w(4x2)"},
  {2, 2, 2, 2, "32x32c/4w(2x2)"},
  {4, 4, 1, 1, "32x32c/1w(4x4)"},
  {4, 4, 2, 1, "64x32c/2w(4x4)"},
};
static const int kNShapes = (int)(sizeof(kShapes) / sizeof(kShapes[0]));

extern "C" int rsg_gram_nshapes(void) { return kNShapes; }
extern "C" const char *rsg_gram_shape_name(int s) { return (s >= 1 && s <= kNShapes) ? kShapes[s - 1].name : "auto"; }
size_t rsg_gram_tile_cols(int s) { const GramShape &g = kShapes[(s >= 1 && s <= kNShapes) ? s - 1 : 0]; return 8 * (size_t)std::max(g.TI * g.WI, g.TJ * g.WJ); }

// Upper-triangular tile list for (L, BI, BJ), cached per process and device.
static const int2 *tile_list(int device, int64_t L, int BI, int BJ, int *ntiles, char *errbuf)
{
  static std::map<std::tuple<int, int64_t, int, int>, std::pair<int2 *, int>> cache;
  auto key = std::make_tuple(device, L, BI, BJ);
  auto it = cache.find(key);
  if (it != cache.end()) { *ntiles = it->second.second; return it->second.first; }
  std::vector<int2> v;
  int64_t nbi = (L + BI - 1) / BI, nbj = (L + BJ - 1) / BJ;
  for (int64_t bi = 0; bi < nbi; bi++)
    for (int64_t bj = 0; bj < nbj; bj++) {
      int64_t jmax = std::min((bj + 1) * BJ, L) - 1;
      if (bi * BI < jmax) v.push_back(make_int2((int)bi, (int)bj));
    }
  int2 *d = nullptr;
  if (hipMalloc(&d, sizeof(int2) * std::max<size_t>(v.size(), 1)) != hipSuccess) { rsg_fail(errbuf, "hipMalloc tile list"); return nullptr; }
  if (!v.empty() && hipMemcpy(d, v.data(), sizeof(int2) * v.size(), hipMemcpyHostToDevice) != hipSuccess) { rsg_fail(errbuf, "tile list copy"); return nullptr; }
  cache[key] = std::make_pair(d, (int)v.size());
  *ntiles = (int)v.size();
  return d;
}

template <int TI, int TJ, int WI, int WJ>
static int launch_i8(rsg_ctx *c, rsg_msa *m, void *out, int kind, char *errbuf)
{
  constexpr int BI = 8 * TI * WI, BJ = 8 * TJ * WJ;
  if (m->Lpad % BI || m->Lpad % BJ || m->Npad % KT) return rsg_fail(errbuf, "gram: padding does not fit the tile");
  int ntiles = 0;
  const int2 *tiles = tile_list(c->device, m->L, BI, BJ, &ntiles, errbuf);
  if (!tiles) return RSG_FAIL;
  if (ntiles == 0) return RSG_OK;
  int per = (ntiles + 7) / 8;
  dim3 grid((unsigned)(8 * per)), block(64 * WI * WJ);
  int t; rsg_tstart(c, "gram_i8", &t);
  if (kind == CNT_U16)
    hipLaunchKernelGGL((k_gram_i8<TI, TJ, WI, WJ, CNT_U16>), grid, block, 0, c->stream, m->d_xt, m->Npad, m->L, out, tiles, ntiles, per);
  else
    hipLaunchKernelGGL((k_gram_i8<TI, TJ, WI, WJ, CNT_U32>), grid, block, 0, c->stream, m->d_xt, m->Npad, m->L, out, tiles, ntiles, per);
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  return RSG_OK;
}

int rsg_gram_counts_i8(rsg_ctx *c, rsg_msa *m, void *out, int kind, char *errbuf)
{
  int s = c->gram_shape;
  if (s < 1 || s > kNShapes) s = (m->L >= 512) ? 1 : 4;   // small alignments: small tiles
  switch (s) {
  case 1: return launch_i8<4, 4, 2, 2>(c, m, out, kind, errbuf);
  case 2: return launch_i8<2, 4, 2, 2>(c, m, out, kind, errbuf);
  case 3: return launch_i8<4, 2, 2, 2>(c, m, out, kind, errbuf);
  case 4: return launch_i8<2, 2, 2, 2>(c, m, out, kind, errbuf);
  case 5: return launch_i8<4, 4, 1, 1>(c, m, out, kind, errbuf);
  case 6: return launch_i8<4, 4, 2, 1>(c, m, out, kind, errbuf);
  }
  return rsg_fail(errbuf, "bad gram shape %d", s);
}

// ------------------------------------------------------------------------------------
// Weighted Gram with split-BF16 weights.
// Operands are expanded once per alignment into feature-major BF16 rows
// (row 4*col + sym, K = sequences contiguous): xw[p] = split part p of w_s where the
// symbol matches, x1 = 1.0 where it matches.
__global__ void k_expand_bf16(const uint8_t *__restrict__ axT, const double *__restrict__ w, int64_t N,
                              int64_t Npad, int64_t Lpad, int nsplit,
                              __hip_bfloat16 *__restrict__ x1, __hip_bfloat16 *__restrict__ xw0,
                              __hip_bfloat16 *__restrict__ xw1, __hip_bfloat16 *__restrict__ xw2)
{
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= Lpad * Npad) return;
  int64_t col = idx / Npad, s = idx - col * Npad;
  uint8_t a = axT[idx];
  double ws = (s < N) ? w[s] : 0.;
  // split: hi = bf16(w), lo = bf16(w - hi), lo2 = bf16(w - hi - lo)
  __hip_bfloat16 p0 = __float2bfloat16((float)ws);
  double r1 = ws - (double)__bfloat162float(p0);
  __hip_bfloat16 p1 = __float2bfloat16((float)r1);
  double r2 = r1 - (double)__bfloat162float(p1);
  __hip_bfloat16 p2 = __float2bfloat16((float)r2);
  __hip_bfloat16 zero = __float2bfloat16(0.f), one = __float2bfloat16(1.f);
  for (int sym = 0; sym < 4; sym++) {
    int64_t o = (col * 4 + sym) * Npad + s;
    bool hit = (a == sym) && (s < N);
    x1[o]  = hit ? one : zero;
    xw0[o] = hit ? p0 : zero;
    xw1[o] = hit ? p1 : zero;
    if (nsplit > 2) xw2[o] = hit ? p2 : zero;
  }
}

int rsg_upload_weights_bf16(rsg_ctx *c, rsg_msa *m, int nsplit, char *errbuf)
{
  if (m->nsplit_ready >= nsplit) return RSG_OK;
  size_t bytes = sizeof(__hip_bfloat16) * 4 * (size_t)m->Lpad * (size_t)m->Npad;
  if (!m->d_x1) HIPCHK(hipMalloc(&m->d_x1, bytes));
  for (int p = 0; p < nsplit; p++) if (!m->d_xw[p]) HIPCHK(hipMalloc(&m->d_xw[p], bytes));
  int64_t n = m->Lpad * m->Npad;
  int t; rsg_tstart(c, "expand_bf16", &t);
  hipLaunchKernelGGL(k_expand_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     m->d_axT, m->d_w, m->N, m->Npad, m->Lpad, nsplit, m->d_x1, m->d_xw[0], m->d_xw[1],
                     nsplit > 2 ? m->d_xw[2] : m->d_xw[1]);
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  m->nsplit_ready = nsplit;
  return RSG_OK;
}

// One wave per 8 x 8 column block (32 x 32 features), WPB waves per work group; operands
// are read straight from HBM/L2 (each 16-byte fragment serves one MFMA).
template <int NS>
__global__ void __launch_bounds__(256)
k_gram_bf16(const __hip_bfloat16 *__restrict__ x1, const __hip_bfloat16 *__restrict__ xw0,
            const __hip_bfloat16 *__restrict__ xw1, const __hip_bfloat16 *__restrict__ xw2,
            int64_t Npad, int64_t L, float *__restrict__ out, const int2 *__restrict__ tiles, int ntiles)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + wave;
  if (t >= ntiles) return;
  const int2 tl = tiles[t];
  const int64_t i0 = (int64_t)tl.x * 8, j0 = (int64_t)tl.y * 8;
  const int khalf = lane >> 5, f = lane & 31;
  const __hip_bfloat16 *pa0 = xw0 + (j0 * 4 + f) * Npad + khalf * 8;
  const __hip_bfloat16 *pa1 = xw1 + (j0 * 4 + f) * Npad + khalf * 8;
  const __hip_bfloat16 *pa2 = xw2 + (j0 * 4 + f) * Npad + khalf * 8;
  const __hip_bfloat16 *pb  = x1  + (i0 * 4 + f) * Npad + khalf * 8;
  v16f acc;
#pragma unroll
  for (int r = 0; r < 16; r++) acc[r] = 0.f;
  for (int64_t k = 0; k < Npad; k += 16) {
    v8bf b  = *(const v8bf *)(pb + k);
    v8bf a0 = *(const v8bf *)(pa0 + k);
    v8bf a1 = *(const v8bf *)(pa1 + k);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b, acc, 0, 0, 0);
    if (NS > 2) {
      v8bf a2 = *(const v8bf *)(pa2 + k);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2, b, acc, 0, 0, 0);
    }
  }
  const int x = lane & 3, ii = (lane & 31) >> 2;
  const int64_t i = i0 + ii;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int64_t j = j0 + 2 * q + khalf;
    if (i < j && j < L)
      *(float4 *)(out + (i * L + j) * 16 + x * 4) = make_float4(acc[4 * q], acc[4 * q + 1], acc[4 * q + 2], acc[4 * q + 3]);
  }
}

int rsg_gram_counts_bf16(rsg_ctx *c, rsg_msa *m, int nsplit, float *out, char *errbuf)
{
  if (rsg_upload_weights_bf16(c, m, nsplit, errbuf) != RSG_OK) return RSG_FAIL;
  int ntiles = 0;
  const int2 *tiles = tile_list(c->device, m->L, 8, 8, &ntiles, errbuf);
  if (!tiles) return RSG_FAIL;
  if (!ntiles) return RSG_OK;
  int t; rsg_tstart(c, "gram_bf16", &t);
  dim3 grid((unsigned)((ntiles + 3) / 4));
  if (nsplit > 2)
    hipLaunchKernelGGL(k_gram_bf16<3>, grid, dim3(256), 0, c->stream, m->d_x1, m->d_xw[0], m->d_xw[1], m->d_xw[2], m->Npad, m->L, out, tiles, ntiles);
  else
    hipLaunchKernelGGL(k_gram_bf16<2>, grid, dim3(256), 0, c->stream, m->d_x1, m->d_xw[0], m->d_xw[1], m->d_xw[1], m->Npad, m->L, out, tiles, ntiles);
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  return RSG_OK;
}
