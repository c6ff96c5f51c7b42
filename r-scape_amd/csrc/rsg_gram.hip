// rsg_gram.hip -- pair-count tables as one-hot Gram contractions on CDNA4 MFMA.
//
// Replaces the O(L^2 N) gather loop of mutual_naive_ppij (src/correlators.c:1695-1780,
// driven by corr_NaivePP :1300-1317).  With GAPASCHAR == 0 (src/rscape_config.h:48) the
// table of a column pair is 4 x 4: c[x][y] = sum_s w_s [a_si == x][a_sj == y] over
// canonical symbols.  Writing X for the N x 4L one-hot matrix, all tables at once are the
// Gram C = X^T diag(w) X; only upper-triangular column blocks are formed.
//
// Staging: one byte per (column, sequence) holding 1 << symbol (0 for gap / N), stored
// K-chunked ("xt": [Npad / 16][Lpad][16 bytes]: the 16 sequences of a chunk for every column,
// chunk after chunk), so the rows of a stage are contiguous in HBM and in LDS and the int8
// Gram moves them with LDS-direct loads.  A wave owns 32 i-columns x 32 j-columns and keeps
// all 16 symbol cells as 16 MFMA accumulators ("symbol-major" tiles): each lane reads
// the 16 staged bytes of its column once per 32 sequences and derives the four symbol
// operands in registers, so LDS traffic is one byte per (column, sequence) per wave.
//
// Unit weights: int8 MFMA (v_mfma_i32_32x32x32_i8), int32 accumulation, bit-exact.
//   The operand for symbol s is the staged byte masked with 0x01 << s, i.e. 2^s or 0;
//   cell (x, y) therefore accumulates 2^(x+y) * count, which the epilogue shifts out.
// Weights: split-BF16 (w = hi + lo [+ lo2]) MFMA (v_mfma_f32_32x32x16_bf16); the i-side
//   operand is 1.0 where the symbol matches, the j-side operand the split weight part;
//   all parts accumulate into one FP32 accumulator.
//
// MFMA 32x32 operand / result lane layout (CDNA3/4): A row m and B column n belong to
// lanes with lane % 32 == m (resp. n); the K values split across lane / 32; result
// register r of lane l holds D[8*(r/4) + 4*(l/32) + r%4][l % 32].  A rows are j columns,
// B columns are i columns.  A Gram uses the same K assignment on both sides, so the
// order in which the hardware pairs K values need not be known.
//
// Output: a 16-cell record per column pair i < j at record rsg_rec(i, j) = j*L + i,
// cell x*4+y (the layout of pp[i][j][IDX(x,y,4)] in the reference); unit-weight records hold
// uint16 counts (N <= 65535), weighted ones FP32.
#include "rsg_device.h"
#include <dlfcn.h>
#include <map>
#include <mutex>
#include <tuple>
#include <algorithm>

typedef int   v4i  __attribute__((ext_vector_type(4)));
typedef int   v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

#define NPAD_Q 256            // sequence padding quantum (covers every K-tile size below)

int64_t rsg_npad(int64_t N) { return ((N + NPAD_Q - 1) / NPAD_Q) * NPAD_Q; }

// ------------------------------------------------------------------------------------
// Pack: ax (N x L, row-major) -> axT (Lpad x Npad raw codes) and xt (K-chunked one-hot
// codes, see above; pack_xt_chunk).  One 64-sequence x 64-column block per work group.
__global__ void k_pack(const uint8_t *__restrict__ ax, int64_t N, int64_t L, int64_t Npad, int64_t Lpad,
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
    axT[c * Npad + s] = tile[tx][r];
  }
  pack_xt_chunk(tile, s0, c0, Lpad, xt);
}

int rsg_pack(rsg_ctx *c, rsg_msa *m, char *errbuf)
{
  dim3 grid((unsigned)(m->Lpad / 64), (unsigned)(m->Npad / 64));
  int t; rsg_tstart(c, "pack", &t);
  hipLaunchKernelGGL(k_pack, grid, dim3(256), 0, c->stream, m->d_ax, m->N, m->L, m->Npad, m->Lpad, m->d_axT, m->d_xt);
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  return RSG_OK;
}

// ------------------------------------------------------------------------------------
// Staging of code bytes for one K-tile in LDS: the j side [chunk (16 seqs)][BJ columns][16 B]
// (CSA bytes per chunk row), then the i side [chunk][BI columns][16 B]; no padding (a
// ds_read_b128 pass of 16 lanes reads 16 consecutive columns: conflict-free).
template <int BI, int BJ, int KT, int NT>
struct CodeStage {
  static constexpr int NKC = KT / 16;
  static constexpr int CSA = BJ * 16, CSB = BI * 16;
  static constexpr int BYTES = NKC * (CSA + CSB);
  static constexpr int NCH = (BI + BJ) * NKC;
  static constexpr int NR = (NCH + NT - 1) / NT;
  static constexpr int NINST = NCH / 64;                    // 64-chunk LDS-direct loads per stage
  // LDS of the int8 Gram: three stage images; with fused marginals at least the epilogue's
  // exchange (8 uint4 per lane and wave, NT * 128 bytes) and row sums (4 doubles per lane of
  // half the waves, NT * 16 bytes)
  static constexpr int lds_bytes(bool pm) { return (pm && NT * 144 > 3 * BYTES) ? NT * 144 : 3 * BYTES; }
};

// VGPR staging (the BF16 Gram): every thread moves exactly NR chunks (NCH is a multiple of
// the thread count for all shapes); consecutive threads take consecutive columns of a chunk
// row, so loads and LDS stores are contiguous.
#define STAGE_GLOAD(St, pre, xt, Lpad, i0, j0, k0, tid)                                        \
  static_assert(St::NCH % NT == 0, "stage chunks must divide evenly among threads");         \
  _Pragma("unroll") for (int r_ = 0; r_ < St::NR; r_++) {                                      \
    const int ch_ = (tid) + r_ * NT;                                                           \
    const int kc_ = ch_ / (BI + BJ), cc_ = ch_ % (BI + BJ);                                    \
    const int64_t col_ = (cc_ < BJ) ? ((j0) + cc_) : ((i0) + cc_ - BJ);                        \
    pre[r_] = *(const uint4 *)((xt) + (((k0) / 16 + kc_) * (Lpad) + col_) * 16);               \
  }
#define STAGE_SSTORE(St, pre, base, tid)                                                        \
  _Pragma("unroll") for (int r_ = 0; r_ < St::NR; r_++) {                                      \
    const int ch_ = (tid) + r_ * NT;                                                           \
    const int kc_ = ch_ / (BI + BJ), cc_ = ch_ % (BI + BJ);                                    \
    uint8_t *p_ = (cc_ < BJ) ? ((base) + kc_ * St::CSA + cc_ * 16)                             \
                             : ((base) + St::NKC * St::CSA + kc_ * St::CSB + (cc_ - BJ) * 16); \
    *(uint4 *)p_ = pre[r_];                                                                    \
  }

// LDS-direct staging (the int8 Gram): global_load_lds_dwordx4 moves 16 bytes per lane from
// the lane's own global address to LDS at M0 + 16 * lane, without VGPRs.  The compiler does
// not model the LDS write of the instruction, so it is issued through inline asm and every
// wait for it is an explicit vmcnt wait.
__device__ __forceinline__ void lds_direct16(const uint8_t *g, uint8_t *l)
{
  const uint32_t la = __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(__attribute__((address_space(3))) uint8_t *)l);
  asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off" :: "s"(la), "v"(g) : "memory", "m0");
}

// One stage (K-tile `kt`) into LDS image `base`: the stage's NCH chunks are the j-side chunk
// rows then the i-side chunk rows, 64 consecutive chunks per wave instruction (a j-side
// image of NKC * BJ chunks is a multiple of 64, so no instruction straddles the sides).
// Every wave issues exactly NINST / NW instructions.
template <class St, int BI, int BJ, int NW>
__device__ __forceinline__ void stage_lds_direct(const uint8_t *__restrict__ xt, int64_t Lpad, int64_t i0, int64_t j0,
                                                 int64_t kt, uint8_t *base, int wave, int lane)
{
  static_assert(St::NINST % NW == 0 && (St::NKC * BJ) % 64 == 0, "LDS-direct stage shape");
#pragma unroll
  for (int q = 0; q < St::NINST / NW; q++) {
    const int c0 = (wave + NW * q) * 64, c = c0 + lane;
    const bool aside = c < St::NKC * BJ;
    const int cs = aside ? c : c - St::NKC * BJ, w = aside ? BJ : BI;
    const int kc = cs / w, cc = cs % w;
    const int64_t colg = (aside ? j0 : i0) + cc;
    lds_direct16(xt + ((kt * St::NKC + kc) * Lpad + colg) * 16, base + c0 * 16);
  }
}

// Cell (x, y) of an accumulator holds count << (x + y); two cells, scale removed, per word.
__device__ __forceinline__ unsigned pack_cells(int a, int sa, int b, int sb)
{
  return ((unsigned)a >> sa) | (((unsigned)b >> sb) << 16);
}

// int8 Gram: 2 x WI x WJ waves; wave w owns the 32 i-columns x 32 j-columns block w % (WI WJ)
// and the symbol rows x of its half (w / (WI WJ)): x in {0, 1} or {2, 3}, 8 accumulators of
// 16 int32 per lane.  Two waves share every SIMD (at most 256 registers each), so one wave's
// LDS and barrier waits are covered by the other's MFMAs.  KT sequences per stage.  The
// epilogue shifts the symbol scale out and writes uint16 (OK = CNT_U16) or uint32 (CNT_U32)
// records; each half writes the cells of its own symbol rows.
// One work-group tile (i0, j0) of one alignment over K-tiles [kt0, kt1); `lds` holds
// St::lds_bytes(pm_i != nullptr).  With part == nullptr the records go to `out`; otherwise the tile's partial
// counts go, unconditionally and in the record format, to partial record jj * BI + ii (ii, jj:
// column offsets in the tile) for a later reduction.
// With pm_i != nullptr (uint16 records over the whole K range only) the epilogue also leaves the
// marginal partials of the tile's 32 x 32 blocks (pm_i / pm_j, layout and order in rsg_device.h).
template <int WI, int WJ, int KT, int OK>
__device__ __forceinline__ void gram_tile_i8(const uint8_t *__restrict__ xt, int64_t Lpad, int64_t L,
                                             void *__restrict__ out, int64_t i0, int64_t j0, uint8_t *lds,
                                             int64_t kt0, int64_t kt1, void *__restrict__ part,
                                             double *__restrict__ pm_i = nullptr, double *__restrict__ pm_j = nullptr,
                                             int64_t nb32 = 0, int64_t tailc = -1)
{
#pragma clang fp contract(off)
  constexpr int BI = 32 * WI, BJ = 32 * WJ, NT = 128 * WI * WJ, KS = KT / 32;
  typedef CodeStage<BI, BJ, KT, NT> St;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int blk = wave % (WI * WJ), half = wave / (WI * WJ), x0 = 2 * half;
  const int wi = blk % WI, wj = blk / WI;
  const int col = lane & 31, khalf = lane >> 5;
  const int offA = (wj * 32 + col) * 16 + khalf * St::CSA;
  const int offB = St::NKC * St::CSA + (wi * 32 + col) * 16 + khalf * St::CSB;

  v16i acc[2][4];                                  // [x - x0][y]
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 4; y++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[x][y][r] = 0;
  // tailc >= 0: the sequences past the last whole K-tile fill one 32-sequence k-step (chunks
  // tailc, tailc + 1; zero codes past N), loaded here straight into the lane's fragments and
  // multiplied after the loop, so the loop runs whole K-tiles only
  uint4 ta = make_uint4(0, 0, 0, 0), tb = make_uint4(0, 0, 0, 0);
  if (tailc >= 0) {
    ta = *(const uint4 *)(xt + ((tailc + khalf) * Lpad + j0 + wj * 32 + col) * 16);
    tb = *(const uint4 *)(xt + ((tailc + khalf) * Lpad + i0 + wi * 32 + col) * 16);
  }

  // three LDS images, loads two stages ahead: stage kt0 + it is read from image it % 3 while
  // stage kt0 + it + 2 lands in image (it + 2) % 3, which was last read in iteration it - 1
  // (ended by a barrier).  Stages past kt1 - 1 reload kt1 - 1 into an image nobody reads.
  constexpr int NW = 2 * WI * WJ, QW = St::NINST / NW; // LDS-direct loads per wave per stage
  static_assert(QW >= 1 && QW <= 15, "vmcnt range");
  const int nit = (int)(kt1 - kt0);
  auto stage_of = [&](int it) -> int64_t { return kt0 + ((it < nit) ? it : nit - 1); };
  stage_lds_direct<St, BI, BJ, NW>(xt, Lpad, i0, j0, stage_of(0), lds, wave, lane);
  stage_lds_direct<St, BI, BJ, NW>(xt, Lpad, i0, j0, stage_of(1), lds + St::BYTES, wave, lane);
  __builtin_amdgcn_s_waitcnt(0x0F70 | QW);     // vmcnt(QW): this wave's loads of stage 0 landed
  __syncthreads();                             // ... and every wave's
  // The wait for stage it + 1 and the barrier sit before the last k-step of stage it: every read
  // of image it % 3 has completed there (the barrier waits for LDS reads), so the loads issued at
  // the next iteration's start may overwrite it, and the first fragments of stage it + 1 are read
  // right after the barrier while the last MFMAs of stage it run.
  static_assert(KS >= 2, "the barrier precedes the last k-step");
  uint4 ra = *(const uint4 *)(lds + offA);
  uint4 rb = *(const uint4 *)(lds + offB);
  for (int it = 0; it < nit; it++) {
    stage_lds_direct<St, BI, BJ, NW>(xt, Lpad, i0, j0, stage_of(it + 2), lds + ((it + 2) % 3) * St::BYTES, wave, lane);
    const uint8_t *base = lds + (it % 3) * St::BYTES;
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      uint4 na = ra, nb = rb;
      if (ks + 1 < KS) {
        na = *(const uint4 *)(base + offA + (ks + 1) * 2 * St::CSA);
        nb = *(const uint4 *)(base + offB + (ks + 1) * 2 * St::CSB);
      } else {
        __builtin_amdgcn_s_waitcnt(0x0F70 | QW); // vmcnt(QW): stage it + 1 landed
        __syncthreads();
        const uint8_t *nbase = lds + ((it + 1) % 3) * St::BYTES;
        na = *(const uint4 *)(nbase + offA);
        nb = *(const uint4 *)(nbase + offB);
      }
      v4i a[4], b[2];
#pragma unroll
      for (int s = 0; s < 4; s++) {
        const unsigned mk = 0x01010101u << s;
        a[s] = (v4i){(int)(ra.x & mk), (int)(ra.y & mk), (int)(ra.z & mk), (int)(ra.w & mk)};
      }
#pragma unroll
      for (int x = 0; x < 2; x++) {
        const unsigned mk = 0x01010101u << (x0 + x);
        b[x] = (v4i){(int)(rb.x & mk), (int)(rb.y & mk), (int)(rb.z & mk), (int)(rb.w & mk)};
      }
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 4; y++)
          acc[x][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[y], b[x], acc[x][y], 0, 0, 0);
      ra = na; rb = nb;
    }
  }
  __builtin_amdgcn_s_waitcnt(0x0F70);          // no LDS-direct load outlives the loop
  if (tailc >= 0) {
    v4i a[4], b[2];
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const unsigned mk = 0x01010101u << s;
      a[s] = (v4i){(int)(ta.x & mk), (int)(ta.y & mk), (int)(ta.z & mk), (int)(ta.w & mk)};
    }
#pragma unroll
    for (int x = 0; x < 2; x++) {
      const unsigned mk = 0x01010101u << (x0 + x);
      b[x] = (v4i){(int)(tb.x & mk), (int)(tb.y & mk), (int)(tb.z & mk), (int)(tb.w & mk)};
    }
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 4; y++)
        acc[x][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[y], b[x], acc[x][y], 0, 0, 0);
  }

  // epilogue: lane owns i column `col`; register r holds j = 8*(r/4) + 4*khalf + r%4.
  // Both destinations are affine in (ii, jj): base + jj * sj + ii * 16 cells.
  const int ii = wi * 32 + col;
  const int64_t i = i0 + ii;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int jj = wj * 32 + 8 * (r >> 2) + 4 * khalf + (r & 3);
    const int64_t j = j0 + jj;
    if (!part && !(i < j && j < L)) continue;
    const int64_t rr = part ? (int64_t)jj * BI + ii : rsg_rec(i, j, L);
    void *const dst = part ? part : out;
    if (OK == CNT_U16) {                        // cells 8 half .. 8 half + 7: the 16 bytes of this half
      uint4 *p = (uint4 *)((uint16_t *)dst + rr * 16) + half;
      *p = make_uint4(pack_cells(acc[0][0][r], x0, acc[0][1][r], x0 + 1), pack_cells(acc[0][2][r], x0 + 2, acc[0][3][r], x0 + 3),
                      pack_cells(acc[1][0][r], x0 + 1, acc[1][1][r], x0 + 2), pack_cells(acc[1][2][r], x0 + 3, acc[1][3][r], x0 + 4));
    } else {
      uint4 *p = (uint4 *)((uint32_t *)dst + rr * 16) + x0;
#pragma unroll
      for (int x = 0; x < 2; x++)
        p[x] = make_uint4((unsigned)acc[x][0][r] >> (x0 + x), (unsigned)acc[x][1][r] >> (x0 + x + 1),
                          (unsigned)acc[x][2][r] >> (x0 + x + 2), (unsigned)acc[x][3][r] >> (x0 + x + 3));
    }
  }
  if (OK != CNT_U16 || pm_i == nullptr) return;

  // marginal partials.  The two halves of a block split its 16 r rows: half h takes
  // r = 8 h .. 8 h + 7 and receives the sibling's cells of those rows (8 x uint16 per pair, 64 KiB
  // for 8 waves) through the LDS images, which every wave has finished with.  No FMA contraction
  // anywhere in this function, so every producer rounds alike.
  __syncthreads();
  uint4 *xch = (uint4 *)lds;
  uint4 own[8];                                   // this half's cells of its own rows, packed alike
#pragma unroll
  for (int rl = 0; rl < 8; rl++) {
    unsigned cs[2][4], co[2][4];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 4; y++) {
        const unsigned lo = (unsigned)acc[x][y][rl] >> (x0 + x + y), hi = (unsigned)acc[x][y][8 + rl] >> (x0 + x + y);
        cs[x][y] = half ? lo : hi;
        co[x][y] = half ? hi : lo;
      }
    xch[(wave * 8 + rl) * 64 + lane] = make_uint4(cs[0][0] | (cs[0][1] << 16), cs[0][2] | (cs[0][3] << 16),
                                                  cs[1][0] | (cs[1][1] << 16), cs[1][2] | (cs[1][3] << 16));
    own[rl] = make_uint4(co[0][0] | (co[0][1] << 16), co[0][2] | (co[0][3] << 16), co[1][0] | (co[1][1] << 16),
                         co[1][2] | (co[1][3] << 16));
  }
  __syncthreads();
  const int sib = blk + (1 - half) * (WI * WJ);
  const int bi32 = (int)(i0 / 32) + wi, bj32 = (int)(j0 / 32) + wj, Li = (int)L, ic = bi32 * 32 + col;
  const bool blkok = bi32 <= bj32 && bj32 < (int)nb32;      // blocks of the pm32 layout only
  double ai[4] = {0., 0., 0., 0.}, cv[32];
#pragma unroll
  for (int rl = 0; rl < 8; rl++) {
    const int r = 8 * half + rl, j = bj32 * 32 + 8 * (r >> 2) + 4 * khalf + (r & 3);
    const uint4 sv = xch[(sib * 8 + rl) * 64 + lane];
    const unsigned sw[4] = {sv.x, sv.y, sv.z, sv.w}, ow[4] = {own[rl].x, own[rl].y, own[rl].z, own[rl].w};
    int colf[4], ro[2] = {0, 0}, rsb[2] = {0, 0}, ns = 0;
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const int o0 = (int)((ow[y >> 1] >> (16 * (y & 1))) & 0xFFFFu), o1 = (int)((ow[2 + (y >> 1)] >> (16 * (y & 1))) & 0xFFFFu);
      const int s0 = (int)((sw[y >> 1] >> (16 * (y & 1))) & 0xFFFFu), s1 = (int)((sw[2 + (y >> 1)] >> (16 * (y & 1))) & 0xFFFFu);
      ro[0] += o0; ro[1] += o1; rsb[0] += s0; rsb[1] += s1;
      colf[y] = o0 + o1 + s0 + s1;
      ns += colf[y];
    }
    const int rs[4] = {half ? rsb[0] : ro[0], half ? rsb[1] : ro[1], half ? ro[0] : rsb[0], half ? ro[1] : rsb[1]};
    const double inv = pm_inv(blkok && ic < j && j < Li, ns);
#pragma unroll
    for (int x = 0; x < 4; x++) ai[x] += pm_term(rs[x], inv);
#pragma unroll
    for (int y = 0; y < 4; y++) cv[4 * rl + y] = pm_term(colf[y], inv);
  }
  const double sj = pm_rscatter32(cv, col);                  // index col: r = 8 half + col / 4, y = col % 4
  const int64_t bq = blkok ? pm_tri(bi32, bj32, nb32) : 0;
  const int rj = 8 * half + (col >> 2), jbo = 8 * (rj >> 2) + 4 * khalf + (rj & 3);
  if (blkok) pm_j[(bq * 32 + jbo) * 4 + (col & 3)] = sj;
  // i side: half 1 hands its sums over r = 8..15 to half 0, which adds them to its own (r = 0..7)
  double *aix = (double *)(lds + NT * 128);
  if (half == 1)
#pragma unroll
    for (int x = 0; x < 4; x++) aix[(blk * 4 + x) * 64 + lane] = ai[x];
  __syncthreads();
  if (half == 0)
#pragma unroll
    for (int x = 0; x < 4; x++) {
      const double si = pm_khalf_sum(ai[x] + aix[(blk * 4 + x) * 64 + lane]);
      if (blkok && khalf == 0) pm_i[(bq * 32 + col) * 4 + x] = si;
    }
}

// K segments of the split tiles: nsu units over the nrem = ntiles - ndp last tiles, tile t
// taking sbase + 1 segments if t < sextra, else sbase (so nsu = nrem * sbase + sextra can match
// the CU count); tile t's units are split_first(t) .. split_first(t) + segments - 1.
__host__ __device__ __forceinline__ int split_first(int t, int sbase, int sextra)
{
  return t * sbase + (t < sextra ? t : sextra);
}
__device__ __forceinline__ void split_unit(int u, int sbase, int sextra, int &t, int &sg, int &ns)
{
  const int big = sextra * (sbase + 1);
  if (u < big) { t = u / (sbase + 1); sg = u % (sbase + 1); ns = sbase + 1; }
  else { t = sextra + (u - big) / sbase; sg = (u - big) % sbase; ns = sbase; }
}

// Grid: first 8 * per_xcd work groups for the first ndp tiles (whole K range), then the nsu
// split units, each over one K segment of one of the last tiles, writing partial records for
// k_gram_fix.  The split keeps the last, partial wave of tiles from leaving most CUs idle; the
// units go to the work-group slots as the full tiles free them, so the spread of the full
// tiles' end times is filled rather than added to the kernel's length.
template <int WI, int WJ, int KT, int OK>
__global__ void __launch_bounds__(128 * WI * WJ) __attribute__((amdgpu_waves_per_eu(2)))
k_gram_i8(const uint8_t *__restrict__ xt, int64_t nkt, int64_t tailc, int64_t Lpad, int64_t L, void *__restrict__ out,
          const int2 *__restrict__ tiles, int ndp, int per_xcd, int ntiles, int sbase, int sextra,
          uint8_t *__restrict__ part, double *__restrict__ pm_part, int *__restrict__ pmflag)
{
  typedef CodeStage<32 * WI, 32 * WJ, KT, 128 * WI * WJ> St;
  __shared__ __attribute__((aligned(16))) uint8_t lds[St::lds_bytes(OK == CNT_U16)];
  const int nsu = (ntiles - ndp) * sbase + sextra;
  int t;
  int64_t k0 = 0, k1 = nkt;
  uint8_t *p = nullptr;
  if ((int)blockIdx.x < 8 * per_xcd) {
    // XCD-aware mapping (work group g runs on XCD g & 7): consecutive tiles of the
    // (bi, bj)-sorted list run on one XCD (one L2)
    const int g = blockIdx.x;
    t = (g & 7) * per_xcd + (g >> 3);
    if (t >= ndp) return;
  } else {
    const int u = blockIdx.x - 8 * per_xcd;
    if (u >= nsu) return;
    int tr, sg, ns;
    split_unit(u, sbase, sextra, tr, sg, ns);
    t = ndp + tr;
    k0 = nkt * sg / ns; k1 = nkt * (sg + 1) / ns;
    p = part + (int64_t)u * (32 * WI) * (32 * WJ) * (OK == CNT_U16 ? 32 : 64);
  }
  if (pmflag && blockIdx.x == 0 && threadIdx.x == 0) *pmflag = 0;    // set by the marginal reduction
  const int2 tl = tiles[t];
  const int64_t nb32 = (L + 31) / 32;
  double *pmi = (pm_part && !p) ? pm_part : nullptr;               // split segments: in k_gram_fix
  gram_tile_i8<WI, WJ, KT, OK>(xt, Lpad, L, out, (int64_t)tl.x * 32 * WI, (int64_t)tl.y * 32 * WJ, lds, k0, k1, p,
                               pmi, pmi ? pmi + pm32_n(nb32) : nullptr, nb32, k1 == nkt ? tailc : -1);
}

// Split tiles: one work group (16 waves) per 32 x 32 block sums the block's partial records over
// the tile's K segments (integer sums: exact; uint16 records add as words, no half exceeding
// 65535) and writes the records; wave w takes row r = w of the MFMA result layout, all its
// segments' loads in flight together.  For
// uint16 records with pm_part the block's marginal partials follow with the arithmetic and order
// of the Gram epilogue: the row and column terms go through LDS, wave 0 adds the row terms over
// r = 0..7 and r = 8..15 and then the two sums, waves 2 and 3 reduce-scatter the column terms of
// r = 0..7 and r = 8..15 (index 4 (r % 8) + y).
#define FIX_WAVES 16
template <int BI, int BJ, int OK>
__global__ void __launch_bounds__(64 * FIX_WAVES)
k_gram_fix(const uint8_t *__restrict__ part, const int2 *__restrict__ tiles, int ndp, int sbase, int sextra, int64_t L,
           void *__restrict__ out, double *__restrict__ pm_part)
{
#pragma clang fp contract(off)
  constexpr int NQ = (OK == CNT_U16) ? 2 : 4;                       // uint4 per record
  constexpr int NBI = BI / 32, NB = NBI * (BJ / 32);
  __shared__ double rowt[16][4][64], colt[16][4][64];
  const int tr = blockIdx.x / NB, sb = blockIdx.x % NB, wi = sb % NBI, wj = sb / NBI;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 31, khalf = lane >> 5;
  const int2 tl = tiles[ndp + tr];
  const int bi32 = tl.x * NBI + wi, bj32 = tl.y * (BJ / 32) + wj, Li = (int)L, ic = bi32 * 32 + col;
  const int ns = sbase + (tr < sextra ? 1 : 0);
  const uint4 *p0 = (const uint4 *)part + (int64_t)split_first(tr, sbase, sextra) * BI * BJ * NQ;
  constexpr int RPW = 16 / FIX_WAVES;                               // rows per wave
  uint4 s[RPW][NQ];
#pragma unroll
  for (int q = 0; q < RPW; q++)
#pragma unroll
    for (int h = 0; h < NQ; h++) s[q][h] = make_uint4(0, 0, 0, 0);
#pragma unroll 4
  for (int k = 0; k < ns; k++)
#pragma unroll
    for (int q = 0; q < RPW; q++) {
      const int r = RPW * wave + q, jb = 8 * (r >> 2) + 4 * khalf + (r & 3);
      const int64_t pr = (int64_t)(wj * 32 + jb) * BI + wi * 32 + col;   // partial record in the tile
#pragma unroll
      for (int h = 0; h < NQ; h++) {
        const uint4 v = p0[((int64_t)k * BI * BJ + pr) * NQ + h];
        s[q][h].x += v.x; s[q][h].y += v.y; s[q][h].z += v.z; s[q][h].w += v.w;
      }
    }
  const int nb32 = (Li + 31) / 32;
  const bool blkok = bi32 <= bj32 && bj32 < nb32;
#pragma unroll
  for (int q = 0; q < RPW; q++) {
    const int r = RPW * wave + q, j = bj32 * 32 + 8 * (r >> 2) + 4 * khalf + (r & 3);
    const bool valid = ic < j && j < Li;
    if (valid)
#pragma unroll
      for (int h = 0; h < NQ; h++) ((uint4 *)out)[rsg_rec(ic, j, L) * NQ + h] = s[q][h];
    if (OK != CNT_U16 || pm_part == nullptr) continue;
    int c[16];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const unsigned w[4] = {s[q][h].x, s[q][h].y, s[q][h].z, s[q][h].w};
#pragma unroll
      for (int e = 0; e < 4; e++) { c[8 * h + 2 * e] = (int)(w[e] & 0xFFFFu); c[8 * h + 2 * e + 1] = (int)(w[e] >> 16); }
    }
    int colf[4] = {0, 0, 0, 0}, rs[4] = {0, 0, 0, 0}, nsp = 0;
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int y = 0; y < 4; y++) { colf[y] += c[x * 4 + y]; rs[x] += c[x * 4 + y]; }
#pragma unroll
    for (int y = 0; y < 4; y++) nsp += colf[y];
    const double inv = pm_inv(blkok && valid, nsp);
#pragma unroll
    for (int x = 0; x < 4; x++) { rowt[r][x][lane] = pm_term(rs[x], inv); colt[r][x][lane] = pm_term(colf[x], inv); }
  }
  if (OK != CNT_U16 || pm_part == nullptr) return;
  __syncthreads();
  if (!blkok || wave == 1 || wave > 3) return;
  double *pm_i = pm_part, *pm_j = pm_part + pm32_n(nb32);
  const int64_t bq = pm_tri(bi32, bj32, nb32);
  if (wave == 0) {
#pragma unroll
    for (int x = 0; x < 4; x++) {
      double a0 = 0., a1 = 0.;
#pragma unroll
      for (int rl = 0; rl < 8; rl++) { a0 += rowt[rl][x][lane]; a1 += rowt[8 + rl][x][lane]; }
      const double si = pm_khalf_sum(a0 + a1);
      if (khalf == 0) pm_i[(bq * 32 + col) * 4 + x] = si;
    }
  } else {
    const int h = wave - 2;
    double cv[32];
#pragma unroll
    for (int rl = 0; rl < 8; rl++)
#pragma unroll
      for (int y = 0; y < 4; y++) cv[4 * rl + y] = colt[8 * h + rl][y][lane];
    const double sj = pm_rscatter32(cv, col);
    const int rj = 8 * h + (col >> 2), jbo = 8 * (rj >> 2) + 4 * khalf + (rj & 3);
    pm_j[(bq * 32 + jbo) * 4 + (col & 3)] = sj;
  }
}

// Batched int8 Gram over many alignments: one work group per (item, tile); tiles are
// pre-sorted by decreasing sequence count (cost).  Large alignments: 64 x 64 columns, 4
// waves per half, 256-sequence K tiles; small ones: 32 x 32 columns, one wave per half,
// 64-sequence K tiles.
template <int WI, int WJ, int KT>
__global__ void __launch_bounds__(128 * WI * WJ) __attribute__((amdgpu_waves_per_eu(2)))
k_gram_i8_batch(const int4 *__restrict__ tiles, const int64_t *__restrict__ desc, const uint8_t *__restrict__ xt,
                uint16_t *__restrict__ cnt, double *__restrict__ pm_part, int pm_field)
{
  typedef CodeStage<32 * WI, 32 * WJ, KT, 128 * WI * WJ> St;
  __shared__ __attribute__((aligned(16))) uint8_t lds[St::lds_bytes(true)];
  const int4 tl = tiles[blockIdx.x];                    // (item, bi, bj, -)
  const int64_t *d = desc + (int64_t)tl.x * RSG_BDESC;  // item descriptor (see rsg_batch.hip)
  const int64_t nb32 = (d[1] + 31) / 32;
  double *pmi = pm_part + d[pm_field];
  gram_tile_i8<WI, WJ, KT, CNT_U16>(xt + d[4], d[3], d[1], cnt + d[5], (int64_t)tl.y * 32 * WI, (int64_t)tl.z * 32 * WJ,
                                    lds, 0, d[2] / KT, nullptr, pmi, pmi + pm32_n(nb32), nb32);
}

// The same tiles in a persistent grid (as many work groups as fit the device at once) fed by
// per-XCD queues: work groups are dispatched to the 8 XCDs round-robin, so group g serves queue
// x = g % 8, which holds the tiles t = x, x + 8, x + 16, ... of the list (each XCD gets the same
// mix of long and short tiles); a pop takes the next index of its queue from an atomic counter.
// The list is sorted by decreasing K length, so this is longest-first greedy scheduling of
// uneven items; eight counters instead of one keep the pops from serialising on one address.
// The counters are cleared by a fill kernel ahead of the launch (graph-safe).
template <int WI, int WJ, int KT>
__global__ void __launch_bounds__(128 * WI * WJ) __attribute__((amdgpu_waves_per_eu(2)))
k_gram_i8_queue(const int4 *__restrict__ tiles, int ntiles, int *__restrict__ queue, const int64_t *__restrict__ desc,
                const uint8_t *__restrict__ xt, uint16_t *__restrict__ cnt, double *__restrict__ pm_part, int pm_field)
{
  typedef CodeStage<32 * WI, 32 * WJ, KT, 128 * WI * WJ> St;
  __shared__ __attribute__((aligned(16))) uint8_t lds[St::lds_bytes(true)];
  __shared__ int next;
  const int x = blockIdx.x & 7;
  int *q = queue + 16 * x;                              // one counter per 64-byte line
  if (threadIdx.x == 0) next = 8 * atomicAdd(q, 1) + x;
  __syncthreads();
  int t = next;
  while (t < ntiles) {
    int nxt = 0;                                        // the next pop overlaps this tile's first loads
    if (threadIdx.x == 0) nxt = 8 * atomicAdd(q, 1) + x;
    const int4 tl = tiles[t];
    const int64_t *d = desc + (int64_t)tl.x * RSG_BDESC;
    const int64_t nb32 = (d[1] + 31) / 32;
    double *pmi = pm_part + d[pm_field];
    gram_tile_i8<WI, WJ, KT, CNT_U16>(xt + d[4], d[3], d[1], cnt + d[5], (int64_t)tl.y * 32 * WI, (int64_t)tl.z * 32 * WJ,
                                      lds, 0, d[2] / KT, nullptr, pmi, pmi + pm32_n(nb32), nb32);
    __syncthreads();                                    // every wave is done with lds and has read next
    if (threadIdx.x == 0) next = nxt;
    __syncthreads();
    t = next;
  }
}

// the 8 queue counters (16 ints apart)
__global__ void k_clear_queues(int *__restrict__ q) { if (threadIdx.x < 128) q[threadIdx.x] = 0; }

int rsg_gram_batch_launch(rsg_ctx *c, const int4 *d_big, int nbig, const int4 *d_small, int nsmall,
                          const int64_t *d_desc, const uint8_t *d_xt, uint16_t *d_cnt, double *d_pm_part, int pm_field,
                          int *d_queue)
{
  const int ncu = c->ncu > 0 ? c->ncu : 256;
  if (c->batch_persistent && d_queue) {
    // Large tiles (64 x 32 columns, tens of microseconds each) keep one work group per tile:
    // the hardware dispatcher is already a longest-first queue over the sorted list, and the
    // queue loop's extra live registers make that tile body spill (measured 30% slower).  The
    // many short 32 x 32 tiles run persistent, saving a dispatch per tile.
    if (nbig > 0)
      hipLaunchKernelGGL((k_gram_i8_batch<2, 1, 256>), dim3((unsigned)nbig), dim3(256), 0, c->stream, d_big, d_desc, d_xt,
                         d_cnt, d_pm_part, pm_field);
    if (nsmall > 0) {                   // 2 waves per group, four groups per CU
      hipLaunchKernelGGL(k_clear_queues, dim3(1), dim3(256), 0, c->stream, d_queue);
      hipLaunchKernelGGL((k_gram_i8_queue<1, 1, 64>), dim3((unsigned)(std::min(nsmall, 4 * ncu) + 7) & ~7u), dim3(128), 0,
                         c->stream, d_small, nsmall, d_queue, d_desc, d_xt, d_cnt, d_pm_part, pm_field);
    }
    return hipGetLastError() == hipSuccess ? RSG_OK : RSG_FAIL;
  }
  if (nbig > 0)
    hipLaunchKernelGGL((k_gram_i8_batch<2, 1, 256>), dim3((unsigned)nbig), dim3(256), 0, c->stream, d_big, d_desc, d_xt, d_cnt,
                       d_pm_part, pm_field);
  if (nsmall > 0)
    hipLaunchKernelGGL((k_gram_i8_batch<1, 1, 64>), dim3((unsigned)nsmall), dim3(128), 0, c->stream, d_small, d_desc, d_xt, d_cnt,
                       d_pm_part, pm_field);
  return hipGetLastError() == hipSuccess ? RSG_OK : RSG_FAIL;
}

// ------------------------------------------------------------------------------------
// Weighted Gram, split-BF16.  Staged like the int8 Gram plus the split weight parts of
// the K-tile's sequences (bf16 bits, [part][KT]).
//
// Operand expansion for symbol s from 4 staged bytes (one dword d): t = (d >> s) & 0x01010101
// holds 0/1 per byte; v_perm spreads bytes 0,1 (or 2,3) into the two 16-bit halves and a
// 24-bit multiply turns each 0/1 half into 0x0000/0xFFFF (mask) or 0x0000/0x3F80 (1.0).
__device__ __forceinline__ void spread(unsigned d, int s, unsigned &lo, unsigned &hi)
{
  unsigned t = (d >> s) & 0x01010101u;
  lo = __builtin_amdgcn_perm(t, t, 0x0C010C00u);
  hi = __builtin_amdgcn_perm(t, t, 0x0C030C02u);
}

template <int WI, int WJ, int KT, int NS>
__global__ void __launch_bounds__(64 * WI * WJ)
k_gram_bf16(const uint8_t *__restrict__ xt, const uint16_t *__restrict__ wsplit, int64_t Npad, int64_t Lpad, int64_t L,
            float *__restrict__ out, const int2 *__restrict__ tiles, int ntiles, int per_xcd)
{
  constexpr int BI = 32 * WI, BJ = 32 * WJ, NT = 64 * WI * WJ, KS = KT / 32;
  typedef CodeStage<BI, BJ, KT, NT> St;
  constexpr int WBYTES = NS * KT * 2;
  constexpr int SB = St::BYTES + WBYTES;
  __shared__ __attribute__((aligned(16))) uint8_t lds[2 * SB];

  const int g = blockIdx.x;
  const int t = (g & 7) * per_xcd + (g >> 3);
  if (t >= ntiles) return;
  const int2 tl = tiles[t];
  const int64_t i0 = (int64_t)tl.x * BI, j0 = (int64_t)tl.y * BJ;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = wave % WI, wj = wave / WI;
  const int col = lane & 31, khalf = lane >> 5;
  const int offA = (wj * 32 + col) * 16 + khalf * St::CSA;
  const int offB = St::NKC * St::CSA + (wi * 32 + col) * 16 + khalf * St::CSB;

  v16f acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[x][y][r] = 0.f;

  uint4 pre[St::NR];
  uint4 wpre = make_uint4(0, 0, 0, 0);            // weight chunk of this thread (if any)
  constexpr int NWCH = WBYTES / 16;               // 16-byte weight chunks per stage
  auto wload = [&](int64_t k0) {
    if (tid < NWCH) {
      int p = tid / (KT / 8), kc = tid % (KT / 8);
      wpre = *(const uint4 *)(wsplit + p * Npad + k0 + kc * 8);
    }
  };
  auto wstore = [&](uint8_t *base) {
    if (tid < NWCH) *(uint4 *)(base + St::BYTES + tid * 16) = wpre;
  };
  const int64_t nkt = Npad / KT;
  STAGE_GLOAD(St, pre, xt, Lpad, i0, j0, 0, tid);
  wload(0);
  STAGE_SSTORE(St, pre, lds, tid);
  wstore(lds);
  __syncthreads();
  for (int64_t kt = 0; kt < nkt; kt++) {
    const int buf = (int)(kt & 1);
    const int64_t knext = ((kt + 1 < nkt) ? kt + 1 : kt) * KT;
    STAGE_GLOAD(St, pre, xt, Lpad, i0, j0, knext, tid);
    wload(knext);
    const uint8_t *base = lds + buf * SB;
    const uint16_t *wl = (const uint16_t *)(base + St::BYTES);
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      const uint4 ra = *(const uint4 *)(base + offA + ks * 2 * St::CSA);
      const uint4 rb = *(const uint4 *)(base + offB + ks * 2 * St::CSB);
#pragma unroll
      for (int h = 0; h < 2; h++) {               // two 16-sequence MFMA steps per 32-sequence chunk
        const unsigned da0 = h ? ra.z : ra.x, da1 = h ? ra.w : ra.y;
        const unsigned db0 = h ? rb.z : rb.x, db1 = h ? rb.w : rb.y;
        uint4 wv[NS];
#pragma unroll
        for (int p = 0; p < NS; p++)
          wv[p] = *(const uint4 *)(wl + p * KT + ks * 32 + khalf * 16 + h * 8);
        v8bf bo[4];
#pragma unroll
        for (int s = 0; s < 4; s++) {
          unsigned l0, h0, l1, h1;
          spread(db0, s, l0, h0); spread(db1, s, l1, h1);
          uint4 v = make_uint4(__umul24(l0, 0x3F80u), __umul24(h0, 0x3F80u),
                               __umul24(l1, 0x3F80u), __umul24(h1, 0x3F80u));
          bo[s] = *(v8bf *)&v;
        }
#pragma unroll
        for (int y = 0; y < 4; y++) {
          unsigned l0, h0, l1, h1;
          spread(da0, y, l0, h0); spread(da1, y, l1, h1);
          const unsigned m0 = __umul24(l0, 0xFFFFu), m1 = __umul24(h0, 0xFFFFu);
          const unsigned m2 = __umul24(l1, 0xFFFFu), m3 = __umul24(h1, 0xFFFFu);
#pragma unroll
          for (int p = 0; p < NS; p++) {
            uint4 v = make_uint4(wv[p].x & m0, wv[p].y & m1, wv[p].z & m2, wv[p].w & m3);
            const v8bf ao = *(v8bf *)&v;
#pragma unroll
            for (int x = 0; x < 4; x++)
              acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ao, bo[x], acc[x][y], 0, 0, 0);
          }
        }
      }
    }
    STAGE_SSTORE(St, pre, lds + (buf ^ 1) * SB, tid);
    wstore(lds + (buf ^ 1) * SB);
    __syncthreads();
  }

  const int64_t i = i0 + wi * 32 + col;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int64_t j = j0 + wj * 32 + 8 * (r >> 2) + 4 * khalf + (r & 3);
    if (i < j && j < L) {
      float4 *p = (float4 *)(out + rsg_rec(i, j, L) * 16);
#pragma unroll
      for (int x = 0; x < 4; x++) p[x] = make_float4(acc[x][0][r], acc[x][1][r], acc[x][2][r], acc[x][3][r]);
    }
  }
}

// ------------------------------------------------------------------------------------
// Tile shapes (the autotuner sweeps these): waves per work group and K-tile.
struct GramShape { int WI, WJ, KT; const char *name; };
static const GramShape kShapes[] = {
  {2, 2, 128, "64x64c/8w/kt128"},
  {2, 2, 64,  "64x64c/8w/kt64"},
  {2, 2, 256, "64x64c/8w/kt256"},
  {1, 1, 128, "32x32c/2w/kt128"},
  {2, 1, 128, "64x32c/4w/kt128"},
  {1, 2, 128, "32x64c/4w/kt128"},
  {2, 1, 256, "64x32c/4w/kt256"},
  {1, 2, 256, "32x64c/4w/kt256"},
  {1, 1, 256, "32x32c/2w/kt256"},
};
static const int kNShapes = (int)(sizeof(kShapes) / sizeof(kShapes[0]));

extern "C" int rsg_gram_nshapes(void) { return kNShapes; }
extern "C" const char *rsg_gram_shape_name(int s) { return (s >= 1 && s <= kNShapes) ? kShapes[s - 1].name : "auto"; }

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

template <int WI, int WJ, int KT, int OK>
static int launch_i8(rsg_ctx *c, rsg_msa *m, void *out, double *pm_part, int *pmflag, char *errbuf)
{
  constexpr int BI = 32 * WI, BJ = 32 * WJ;
  if (m->Lpad % BI || m->Lpad % BJ || m->Npad % KT) return rsg_fail(errbuf, "gram: padding does not fit the tile");
  int ntiles = 0;
  const int2 *tiles;
  if (c->tiles_fixed) {                      // an allocation-free call brings its own tile list
    if (BI != 32 || BJ != 32) return rsg_fail(errbuf, "gram: a fixed tile list needs 32 x 32-column tiles");
    tiles = c->tiles_fixed; ntiles = c->ntiles_fixed;
  } else tiles = tile_list(c->device, m->L, BI, BJ, &ntiles, errbuf);
  if (!tiles) return RSG_FAIL;
  if (ntiles == 0) return RSG_OK;
  if (OK == CNT_U16 && m->N > RSG_U16_MAXN) return rsg_fail(errbuf, "gram: uint16 records hold up to %d sequences", RSG_U16_MAXN);
  // whole waves of tiles run over the full K range; when there are few waves (under four
  // tiles per work-group slot) and the last wave is at most half full, its tiles are split along
  // K into one segment per CU, which run after the full tiles in the slots the first of them
  // free (the two work groups that share a CU finish far apart), then reduced (with many waves
  // the tail is small and the partial traffic costs more).  A CU holds 8 waves of this kernel
  // (two per SIMD): 4 / (WI WJ) work groups.
  const int ncu = c->ncu > 0 ? c->ncu : 256;
  const int slots = std::max(ncu * 4 / (WI * WJ), 1);
  // whole K-tiles; when the sequences past the last whole one fit one 32-sequence k-step, that
  // step runs from registers after the loop instead of as a K-tile of mostly padding
  int64_t nkt = m->Npad / KT, tailc = -1;
  {
    const int64_t nfull = m->N / KT, r = m->N - nfull * KT;
    if (nfull >= 1 && r > 0 && r <= 32) { nkt = nfull; tailc = nfull * (KT / 16); }
  }
  int ndp = ntiles, sbase = 1, sextra = 0;
  const int rem = ntiles % slots;
  if (c->gram_split && ntiles < 4 * slots && rem > 0 && rem <= slots / 2 && nkt >= 2) {
    ndp = ntiles - rem;
    const int nsu = (int)std::min<int64_t>((int64_t)ncu, (int64_t)rem * nkt);
    sbase = nsu / rem; sextra = nsu % rem;
  }
  const int nrem = ntiles - ndp, nsu = nrem * sbase + sextra;
  uint8_t *part = nullptr;
  if (nrem > 0) {
    const size_t need = (size_t)nsu * BI * BJ * rsg_rec_bytes(OK);
    if (c->gscr_cap < need && c->no_alloc) return rsg_fail(errbuf, "gram: split-K partials need an allocation");
    if (c->gscr_cap < need) {
      if (c->d_gscr) hipFree(c->d_gscr);
      c->d_gscr = nullptr; c->gscr_cap = 0;
      HIPCHK(hipMalloc(&c->d_gscr, need));
      c->gscr_cap = need;
    }
    part = (uint8_t *)c->d_gscr;
  }
  const int per = (ndp + 7) / 8;
  dim3 grid((unsigned)(8 * per + nsu)), block(128 * WI * WJ);
  // one timer over the Gram and its split-K fix-up
  int t; rsg_tstart(c, "gram_i8", &t);
  hipLaunchKernelGGL((k_gram_i8<WI, WJ, KT, OK>), grid, block, 0, c->stream, m->d_xt, nkt, tailc, m->Lpad, m->L, out,
                     tiles, ndp, per, ntiles, sbase, sextra, part, pm_part, pmflag);
  if (nrem > 0)
    hipLaunchKernelGGL((k_gram_fix<BI, BJ, OK>), dim3((unsigned)(nrem * (BI / 32) * (BJ / 32))), dim3(64 * FIX_WAVES), 0, c->stream,
                       part, tiles, ndp, sbase, sextra, m->L, out, pm_part);
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  return RSG_OK;
}

// The autotuner's table: the fastest shape per size bucket (ceil(log2 N), ceil(log2 L)), one
// "logN logL shape" line each ('#' starts a comment).  Loaded once per process from
// gram_tiles.tsv next to the library (or rsg_gram_tiles_load); sizes without an entry use
// the built-in rule below.
static std::mutex g_tiles_mu;
static std::map<std::pair<int, int>, int> g_tiles;
static bool g_tiles_tried = false;

static int ceil_log2(int64_t v)
{
  int k = 0;
  while (((int64_t)1 << k) < v) k++;
  return k;
}

static int parse_tiles(const char *path, std::map<std::pair<int, int>, int> &t, char *errbuf)
{
  FILE *f = path ? fopen(path, "r") : nullptr;
  if (!f) return rsg_fail(errbuf, "rsg_gram_tiles_load: cannot open %s", path ? path : "(null)");
  char line[256];
  int ln = 0;
  while (fgets(line, sizeof line, f)) {
    ln++;
    char *h = strchr(line, '#');
    if (h) *h = 0;
    int a, b, sh;
    char extra;
    const int n = sscanf(line, "%d %d %d %c", &a, &b, &sh, &extra);
    if (n <= 0) continue;                               // blank or comment
    if (n != 3 || a < 0 || a > 40 || b < 0 || b > 40 || sh < 1 || sh > kNShapes) {
      fclose(f);
      return rsg_fail(errbuf, "rsg_gram_tiles_load: %s line %d: expected \"logN logL shape\" with shape 1..%d", path, ln,
                      kNShapes);
    }
    t[std::make_pair(a, b)] = sh;
  }
  fclose(f);
  return RSG_OK;
}

extern "C" int rsg_gram_tiles_load(const char *path, char *errbuf)
{
  std::map<std::pair<int, int>, int> t;
  if (parse_tiles(path, t, errbuf) != RSG_OK) return RSG_FAIL;
  std::lock_guard<std::mutex> g(g_tiles_mu);
  g_tiles.swap(t);
  g_tiles_tried = true;
  return RSG_OK;
}

static int table_shape(int64_t N, int64_t L)
{
  std::lock_guard<std::mutex> g(g_tiles_mu);
  if (!g_tiles_tried) {
    g_tiles_tried = true;
    Dl_info info;
    if (dladdr((void *)&rsg_gram_tiles_load, &info) && info.dli_fname) {
      std::string p(info.dli_fname);
      const size_t k = p.rfind('/');
      p = (k == std::string::npos ? std::string(".") : p.substr(0, k)) + "/gram_tiles.tsv";
      std::map<std::pair<int, int>, int> t;
      char eb[256];
      if (FILE *f = fopen(p.c_str(), "r")) {
        fclose(f);
        if (parse_tiles(p.c_str(), t, eb) == RSG_OK) g_tiles.swap(t);   // a malformed table is ignored
      }
    }
  }
  auto it = g_tiles.find(std::make_pair(ceil_log2(N), ceil_log2(L)));
  return it == g_tiles.end() ? 0 : it->second;
}

static int pick_shape(const rsg_ctx *c, const rsg_msa *m)
{
  int s = c->gram_shape;
  if (s >= 1 && s <= kNShapes) return s;
  s = table_shape(m->N, m->L);
  if (s >= 1 && s <= kNShapes) return s;
  // autotuned on MI355X: 64 x 32-column tiles of four waves with 256-sequence K tiles win or tie
  // from N = 2000 to 20000 and L = 500 to 5000 (two work groups per CU, so one's epilogue runs
  // under the other's K loop); below ~1000 sequences (two or three K tiles) 32 x 32 tiles of two
  // waves are as fast or faster
  return (m->N < 1024) ? 4 : 7;
}

int rsg_gram_counts_i8(rsg_ctx *c, rsg_msa *m, void *out, int kind, double *pm_part, int *pmflag, char *errbuf)
{
  if (kind == CNT_U32)            // more than 65535 sequences: the two default shapes only, no fused marginals
    return (m->L >= 768) ? launch_i8<2, 2, 256, CNT_U32>(c, m, out, nullptr, nullptr, errbuf)
                         : launch_i8<1, 1, 128, CNT_U32>(c, m, out, nullptr, nullptr, errbuf);
  if (kind != CNT_U16) return rsg_fail(errbuf, "gram: bad record kind %d", kind);
  switch (pick_shape(c, m)) {
  case 1: return launch_i8<2, 2, 128, CNT_U16>(c, m, out, pm_part, pmflag, errbuf);
  case 2: return launch_i8<2, 2, 64, CNT_U16>(c, m, out, pm_part, pmflag, errbuf);
  case 3: return launch_i8<2, 2, 256, CNT_U16>(c, m, out, pm_part, pmflag, errbuf);
  case 4: return launch_i8<1, 1, 128, CNT_U16>(c, m, out, pm_part, pmflag, errbuf);
  case 5: return launch_i8<2, 1, 128, CNT_U16>(c, m, out, pm_part, pmflag, errbuf);
  case 6: return launch_i8<1, 2, 128, CNT_U16>(c, m, out, pm_part, pmflag, errbuf);
  case 7: return launch_i8<2, 1, 256, CNT_U16>(c, m, out, pm_part, pmflag, errbuf);
  case 8: return launch_i8<1, 2, 256, CNT_U16>(c, m, out, pm_part, pmflag, errbuf);
  case 9: return launch_i8<1, 1, 256, CNT_U16>(c, m, out, pm_part, pmflag, errbuf);
  }
  return rsg_fail(errbuf, "bad gram shape");
}

// split weights: part p of w_s as bf16 bits, [NS][Npad]; hi = bf16(w), lo = bf16(w - hi), ...
__global__ void k_split_weights(const double *__restrict__ w, int64_t N, int64_t Npad, int nsplit,
                                uint16_t *__restrict__ out)
{
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= Npad) return;
  double r = (s < N) ? w[s] : 0.;
  for (int p = 0; p < nsplit; p++) {
    __hip_bfloat16 b = __float2bfloat16((float)r);
    r -= (double)__bfloat162float(b);
    out[p * Npad + s] = __bfloat16_as_ushort(b);
  }
}

int rsg_upload_weights_bf16(rsg_ctx *c, rsg_msa *m, int nsplit, char *errbuf)
{
  if (m->nsplit_ready == nsplit) return RSG_OK;
  if (!m->d_wsplit) HIPCHK(hipMalloc(&m->d_wsplit, sizeof(uint16_t) * 3 * (size_t)m->Npad));
  hipLaunchKernelGGL(k_split_weights, dim3((unsigned)((m->Npad + 255) / 256)), dim3(256), 0, c->stream,
                     m->d_w, m->N, m->Npad, nsplit, m->d_wsplit);
  HIPCHK(hipGetLastError());
  m->nsplit_ready = nsplit;
  return RSG_OK;
}

template <int WI, int WJ, int KT, int NS>
static int launch_bf16(rsg_ctx *c, rsg_msa *m, float *out, char *errbuf)
{
  constexpr int BI = 32 * WI, BJ = 32 * WJ;
  if (m->Lpad % BI || m->Lpad % BJ || m->Npad % KT) return rsg_fail(errbuf, "gram: padding does not fit the tile");
  int ntiles = 0;
  const int2 *tiles;
  if (c->tiles_fixed) {                      // an allocation-free call brings its own tile list
    if (BI != 32 || BJ != 32) return rsg_fail(errbuf, "gram: a fixed tile list needs 32 x 32-column tiles");
    tiles = c->tiles_fixed; ntiles = c->ntiles_fixed;
  } else tiles = tile_list(c->device, m->L, BI, BJ, &ntiles, errbuf);
  if (!tiles) return RSG_FAIL;
  if (ntiles == 0) return RSG_OK;
  int per = (ntiles + 7) / 8;
  int t; rsg_tstart(c, "gram_bf16", &t);
  hipLaunchKernelGGL((k_gram_bf16<WI, WJ, KT, NS>), dim3((unsigned)(8 * per)), dim3(64 * WI * WJ), 0, c->stream,
                     m->d_xt, m->d_wsplit, m->Npad, m->Lpad, m->L, out, tiles, ntiles, per);
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  return RSG_OK;
}

int rsg_gram_counts_bf16(rsg_ctx *c, rsg_msa *m, int nsplit, float *out, char *errbuf)
{
  if (rsg_upload_weights_bf16(c, m, nsplit, errbuf) != RSG_OK) return RSG_FAIL;
  const bool big = m->L >= 768;
  if (nsplit > 2) return big ? launch_bf16<2, 2, 64, 3>(c, m, out, errbuf) : launch_bf16<1, 1, 64, 3>(c, m, out, errbuf);
  return big ? launch_bf16<2, 2, 64, 2>(c, m, out, errbuf) : launch_bf16<1, 1, 64, 2>(c, m, out, errbuf);
}

// ------------------------------------------------------------------------------------
// Weighted Gram from int8 weight limbs (exact integer accumulation).  Each sequence weight is
// rounded to 21-bit fixed point relative to the largest weight, q_s = round(w_s / W) with
// W = max_s w_s / (2^21 - 1), and split into three 7-bit limbs q = q0 + 2^7 q1 + 2^14 q2.  The
// int8 MFMA then forms, per limb k, G_k[x][y] = sum_s q_{s,k} [a_si = x][a_sj = y] exactly in
// int32 (the j-side operand is the one-hot code masked for symbol y, scaled by 2^y; the i-side
// operand is the limb byte where the code is symbol x, else 0), and the epilogue combines
// W (G_0 + 2^7 G_1 + 2^14 G_2) in FP64 (exact) before the FP32 record store.  The only error
// is the weight rounding: |w_s - W q_s| <= W / 2 = 2^-22 max_s w_s.
//
// Layout of the limb panel: [Npad / 16][4][16] bytes, limbs 0..2 of the 16 sequences of a
// K-chunk (slot 3 zero), so one stage's limbs are NKC * 64 contiguous bytes, staged into LDS
// behind the stage's codes by wave 0 (one more LDS-direct instruction for that wave).
#define W8_BITS 21
__global__ void __launch_bounds__(1024) k_quant_weights(const double *__restrict__ w, int64_t N, int64_t Npad,
                                                         uint8_t *__restrict__ wq, double *__restrict__ scale)
{
  __shared__ double red[1024];
  const int tid = threadIdx.x;
  double mx = 0.;
  for (int64_t s = tid; s < N; s += 1024) mx = fmax(mx, w[s]);
  red[tid] = mx;
  __syncthreads();
  for (int st = 512; st > 0; st >>= 1) {
    if (tid < st) red[tid] = fmax(red[tid], red[tid + st]);
    __syncthreads();
  }
  const double wmax = red[0];
  const double qmax = (double)((1 << W8_BITS) - 1);
  const double W = (wmax > 0.) ? wmax / qmax : 1.;
  if (tid == 0) *scale = W;
  for (int64_t s = tid; s < Npad; s += 1024) {
    double qd = (s < N && wmax > 0.) ? rint(w[s] / W) : 0.;
    qd = fmin(fmax(qd, 0.), qmax);
    const uint32_t q = (uint32_t)qd;
    uint8_t *p = wq + (s / 16) * 64 + (s % 16);
    p[0] = (uint8_t)(q & 127u); p[16] = (uint8_t)((q >> 7) & 127u); p[32] = (uint8_t)((q >> 14) & 127u); p[48] = 0;
  }
}

// One 32 x 32-column tile, four waves: wave x owns the i-side symbol x and all four j-side
// symbols for the three limbs (12 accumulators of 16 int32 per lane).  Staging as in the int8
// Gram (three LDS images, LDS-direct loads two stages ahead), each image followed by the
// stage's limb bytes.
// K-tiles [kt0, kt1).  With SPLIT (a split-K segment) the combined limb sums
// g = G_0 + 2^7 G_1 + 2^14 G_2 of every cell of every pair of the tile go, unconditionally, to
// the int64 partial record jj * 32 + ii (16 cells) for k_gram_w8_fix, and no record or nseff is
// written.
// (SPLIT: the partial-sum epilogue, a template argument so that each variant of the loop gets its
// own register allocation -- with both epilogues after one loop the kernel spilled.)
template <int KT, bool SPLIT>
__device__ __forceinline__ void gram_tile_w8(const uint8_t *__restrict__ xt, const uint8_t *__restrict__ wq, double W,
                                             int64_t Lpad, int64_t L, float *__restrict__ out, double *__restrict__ nsx,
                                             int64_t i0, int64_t j0, uint8_t *lds, int64_t kt0, int64_t kt1,
                                             int64_t *__restrict__ part)
{
#pragma clang fp contract(off)
  constexpr int NT = 256, KS = KT / 32, NW = 4;
  typedef CodeStage<32, 32, KT, NT> St;
  constexpr int IS = St::BYTES + St::NKC * 64;     // one image: codes, then limbs
  constexpr int WCH = St::NKC * 4;                 // 16-byte limb chunks per stage
  constexpr int QW = St::NINST / NW;               // code loads per wave per stage (+1 for wave 0)
  static_assert(QW >= 1 && QW + 1 <= 15 && WCH <= 64, "stage shape");
  static_assert(KS >= 2, "the barrier precedes the last k-step");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, x = wave;
  const int col = lane & 31, khalf = lane >> 5;
  const int offA = col * 16 + khalf * St::CSA;                           // j side
  const int offB = St::NKC * St::CSA + col * 16 + khalf * St::CSB;       // i side
  const int offW = St::BYTES + khalf * 64;                               // limbs of the lane's chunk

  v16i acc[3][4];                                  // [limb][y]
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int y = 0; y < 4; y++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[k][y][r] = 0;

  auto stage = [&](int64_t kt, uint8_t *base) {
    stage_lds_direct<St, 32, 32, NW>(xt, Lpad, i0, j0, kt, base, wave, lane);
    if (wave == 0 && lane < WCH) lds_direct16(wq + (kt * WCH + lane) * 16, base + St::BYTES);
  };
  auto wait_stage = [&]() {                        // this wave's loads of the older stage landed
    if (wave == 0) __builtin_amdgcn_s_waitcnt(0x0F70 | (QW + 1));
    else           __builtin_amdgcn_s_waitcnt(0x0F70 | QW);
  };
  const int nit = (int)(kt1 - kt0);
  auto stage_of = [&](int it) -> int64_t { return kt0 + ((it < nit) ? it : nit - 1); };
  stage(stage_of(0), lds);
  stage(stage_of(1), lds + IS);
  wait_stage();
  __syncthreads();
  uint4 ra = *(const uint4 *)(lds + offA);
  uint4 rb = *(const uint4 *)(lds + offB);
  for (int it = 0; it < nit; it++) {
    stage(stage_of(it + 2), lds + ((it + 2) % 3) * IS);
    const uint8_t *base = lds + (it % 3) * IS;
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      // this k-step's limbs: read before the last k-step's barrier, after which the image may
      // be overwritten by the next iteration's loads
      const uint4 *wp = (const uint4 *)(base + offW + ks * 128);
      const uint4 w0 = wp[0], w1 = wp[1], w2 = wp[2];
      uint4 na = ra, nb = rb;
      if (ks + 1 < KS) {
        na = *(const uint4 *)(base + offA + (ks + 1) * 2 * St::CSA);
        nb = *(const uint4 *)(base + offB + (ks + 1) * 2 * St::CSB);
      } else {
        wait_stage();
        __syncthreads();
        const uint8_t *nbase = lds + ((it + 1) % 3) * IS;
        na = *(const uint4 *)(nbase + offA);
        nb = *(const uint4 *)(nbase + offB);
      }
      // i side: byte 0xFF where the code is symbol x, else 0
      const unsigned t0 = (rb.x >> x) & 0x01010101u, t1 = (rb.y >> x) & 0x01010101u,
                     t2 = (rb.z >> x) & 0x01010101u, t3 = (rb.w >> x) & 0x01010101u;
      const unsigned m0 = (t0 << 8) - t0, m1 = (t1 << 8) - t1, m2 = (t2 << 8) - t2, m3 = (t3 << 8) - t3;
      v4i a[4];
#pragma unroll
      for (int y = 0; y < 4; y++) {
        const unsigned mk = 0x01010101u << y;
        a[y] = (v4i){(int)(ra.x & mk), (int)(ra.y & mk), (int)(ra.z & mk), (int)(ra.w & mk)};
      }
      const v4i b0 = (v4i){(int)(m0 & w0.x), (int)(m1 & w0.y), (int)(m2 & w0.z), (int)(m3 & w0.w)};
#pragma unroll
      for (int y = 0; y < 4; y++) acc[0][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[y], b0, acc[0][y], 0, 0, 0);
      const v4i b1 = (v4i){(int)(m0 & w1.x), (int)(m1 & w1.y), (int)(m2 & w1.z), (int)(m3 & w1.w)};
#pragma unroll
      for (int y = 0; y < 4; y++) acc[1][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[y], b1, acc[1][y], 0, 0, 0);
      const v4i b2 = (v4i){(int)(m0 & w2.x), (int)(m1 & w2.y), (int)(m2 & w2.z), (int)(m3 & w2.w)};
#pragma unroll
      for (int y = 0; y < 4; y++) acc[2][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[y], b2, acc[2][y], 0, 0, 0);
      ra = na; rb = nb;
    }
  }
  __builtin_amdgcn_s_waitcnt(0x0F70);             // no LDS-direct load outlives the loop

  if constexpr (SPLIT) {                           // a split-K segment: the exact partial sums
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int jj = 8 * (r >> 2) + 4 * khalf + (r & 3);
      int64_t g[4];
#pragma unroll
      for (int y = 0; y < 4; y++)
        g[y] = (int64_t)((unsigned)acc[0][y][r] >> y) + 128 * (int64_t)((unsigned)acc[1][y][r] >> y) +
               16384 * (int64_t)((unsigned)acc[2][y][r] >> y);
      longlong2 *p = (longlong2 *)(part + ((int64_t)jj * 32 + col) * 16 + x * 4);
      p[0] = make_longlong2(g[0], g[1]);
      p[1] = make_longlong2(g[2], g[3]);
    }
    return;
  }

  // epilogue: lane owns i column `col`; register r holds j = 8*(r/4) + 4*khalf + r%4; cells
  // x*4 .. x*4+3 of each record are this wave's.  nseff of every pair comes from the exact
  // integer sums (the FP32 cells each carry a rounding, and ngap = total - nseff cancels most of
  // nseff): the four waves' row sums meet in LDS, which every wave has finished with after the
  // barrier, and one rounding W * sum goes to nsx.
  __syncthreads();
  int64_t *xs = (int64_t *)lds;                    // [wave][r][lane]
  const int64_t i = i0 + col;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int64_t j = j0 + 8 * (r >> 2) + 4 * khalf + (r & 3);
    int64_t g[4], sx = 0;
#pragma unroll
    for (int y = 0; y < 4; y++) {
      g[y] = (int64_t)((unsigned)acc[0][y][r] >> y) + 128 * (int64_t)((unsigned)acc[1][y][r] >> y) +
             16384 * (int64_t)((unsigned)acc[2][y][r] >> y);
      sx += g[y];
    }
    xs[(wave * 16 + r) * 64 + lane] = sx;
    if (i < j && j < L)
      *(float4 *)(out + rsg_rec(i, j, L) * 16 + x * 4) =
          make_float4((float)((double)g[0] * W), (float)((double)g[1] * W), (float)((double)g[2] * W), (float)((double)g[3] * W));
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int r = 4 * q + wave;
    const int64_t j = j0 + 8 * (r >> 2) + 4 * khalf + (r & 3);
    if (!(i < j && j < L)) continue;
    const int64_t tot = xs[(0 * 16 + r) * 64 + lane] + xs[(1 * 16 + r) * 64 + lane] + xs[(2 * 16 + r) * 64 + lane] +
                        xs[(3 * 16 + r) * 64 + lane];
    nsx[rsg_rec(i, j, L)] = (double)tot * W;
  }
}

// Grid: 8 * per_xcd work groups for the first ndp tiles (whole K range, neighbouring tiles on one
// XCD), then the split units of the last tiles (K segments as in k_gram_i8: tile ndp + t takes
// sbase + 1 segments if t < sextra, else sbase), which run in the slots the full tiles free.
template <int KT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_gram_w8(const uint8_t *__restrict__ xt, const uint8_t *__restrict__ wq, const double *__restrict__ wscale, int64_t Npad,
          int64_t Lpad, int64_t L, float *__restrict__ out, double *__restrict__ nsx, const int2 *__restrict__ tiles,
          int ndp, int per_xcd, int ntiles, int sbase, int sextra, int64_t *__restrict__ part)
{
  typedef CodeStage<32, 32, KT, 256> St;
  __shared__ __attribute__((aligned(16))) uint8_t lds[3 * (St::BYTES + St::NKC * 64)];
  const int64_t nkt = Npad / KT;
  const int g = blockIdx.x;
  int t;
  int64_t k0 = 0, k1 = nkt;
  int64_t *p = nullptr;
  if (g < 8 * per_xcd) {
    t = (g & 7) * per_xcd + (g >> 3);
    if (t >= ndp) return;
  } else {
    const int u = g - 8 * per_xcd, nsu = (ntiles - ndp) * sbase + sextra;
    if (u >= nsu) return;
    int tr, sg, ns;
    split_unit(u, sbase, sextra, tr, sg, ns);
    t = ndp + tr;
    k0 = nkt * sg / ns; k1 = nkt * (sg + 1) / ns;
    p = part + (int64_t)u * 1024 * 16;
  }
  const int2 tl = tiles[t];
  if (p) gram_tile_w8<KT, true>(xt, wq, *wscale, Lpad, L, out, nsx, (int64_t)tl.x * 32, (int64_t)tl.y * 32, lds, k0, k1, p);
  else   gram_tile_w8<KT, false>(xt, wq, *wscale, Lpad, L, out, nsx, (int64_t)tl.x * 32, (int64_t)tl.y * 32, lds, k0, k1, p);
}

// Split tiles: four work groups (4 waves each) per tile add the tile's int64 partial sums over
// its K segments (exact) and write the records and nseff with the arithmetic of the
// gram_tile_w8 epilogue (one FP64 product and one FP32 rounding per cell; nseff from the exact
// total), so they equal the unsplit kernel's bit for bit.  Wave w of group q takes row
// r = 4 q + w of the MFMA layout.
__global__ void __launch_bounds__(256)
k_gram_w8_fix(const int64_t *__restrict__ part, const int2 *__restrict__ tiles, int ndp, int sbase, int sextra,
              const double *__restrict__ wscale, int64_t L, float *__restrict__ out, double *__restrict__ nsx)
{
#pragma clang fp contract(off)
  const int tr = blockIdx.x >> 2, wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 31, khalf = lane >> 5;
  const int2 tl = tiles[ndp + tr];
  const int ns = sbase + (tr < sextra ? 1 : 0);
  const double W = *wscale;
  const int r = 4 * (blockIdx.x & 3) + wave, jj = 8 * (r >> 2) + 4 * khalf + (r & 3);
  const int64_t i = (int64_t)tl.x * 32 + col, j = (int64_t)tl.y * 32 + jj;
  if (!(i < j && j < L)) return;
  const longlong2 *p0 = (const longlong2 *)(part + (int64_t)split_first(tr, sbase, sextra) * 1024 * 16 + ((int64_t)jj * 32 + col) * 16);
  int64_t g[16];
#pragma unroll
  for (int c = 0; c < 16; c++) g[c] = 0;
#pragma unroll 4
  for (int k = 0; k < ns; k++) {
    const longlong2 *q = p0 + (int64_t)k * 1024 * 8;
#pragma unroll
    for (int h = 0; h < 8; h++) { const longlong2 v = q[h]; g[2 * h] += v.x; g[2 * h + 1] += v.y; }
  }
  int64_t tot = 0;
  float *o = out + rsg_rec(i, j, L) * 16;
#pragma unroll
  for (int x = 0; x < 4; x++) {
    int64_t sx = 0;
#pragma unroll
    for (int y = 0; y < 4; y++) sx += g[x * 4 + y];
    tot += sx;
    *(float4 *)(o + x * 4) = make_float4((float)((double)g[x * 4] * W), (float)((double)g[x * 4 + 1] * W),
                                         (float)((double)g[x * 4 + 2] * W), (float)((double)g[x * 4 + 3] * W));
  }
  nsx[rsg_rec(i, j, L)] = (double)tot * W;
}

int rsg_gram_counts_w8(rsg_ctx *c, rsg_msa *m, float *out, char *errbuf)
{
  constexpr int KT = 256;
  if (m->Npad % KT || m->Lpad % 32) return rsg_fail(errbuf, "gram: padding does not fit the tile");
  if (m->N > (1 << 21)) return rsg_fail(errbuf, "weighted int8 counts: at most %d sequences", 1 << 21);
  if (!m->d_wq) {
    HIPCHK(hipMalloc(&m->d_wq, (size_t)m->Npad * 4 + 64));
  }
  double *d_scale = (double *)(m->d_wq + (size_t)m->Npad * 4);
  if (!m->wq_ready) {
    hipLaunchKernelGGL(k_quant_weights, dim3(1), dim3(1024), 0, c->stream, m->d_w, m->N, m->Npad, m->d_wq, d_scale);
    HIPCHK(hipGetLastError());
    m->wq_ready = 1;
  }
  if (!m->d_nsx) HIPCHK(hipMalloc(&m->d_nsx, sizeof(double) * (size_t)m->L * (size_t)m->L));
  int ntiles = 0;
  const int2 *tiles;
  if (c->tiles_fixed) { tiles = c->tiles_fixed; ntiles = c->ntiles_fixed; }
  else tiles = tile_list(c->device, m->L, 32, 32, &ntiles, errbuf);
  if (!tiles) return RSG_FAIL;
  if (ntiles == 0) return RSG_OK;
  // the tail as in launch_i8: a CU holds two of these 4-wave work groups; when the last wave of
  // tiles is at most half full its tiles are split along K, one segment per CU, after the full
  // tiles, and k_gram_w8_fix adds the exact partial sums
  const int ncu = c->ncu > 0 ? c->ncu : 256, slots = 2 * ncu;
  const int64_t nkt = m->Npad / KT;
  int ndp = ntiles, sbase = 1, sextra = 0;
  const int rem = ntiles % slots;
  if (c->gram_split && ntiles < 4 * slots && rem > 0 && rem <= slots / 2 && nkt >= 2) {
    ndp = ntiles - rem;
    const int nsu = (int)std::min<int64_t>((int64_t)ncu, (int64_t)rem * nkt);
    sbase = nsu / rem; sextra = nsu % rem;
  }
  const int nrem = ntiles - ndp, nsu = nrem * sbase + sextra;
  int64_t *part = nullptr;
  if (nrem > 0) {
    const size_t need = (size_t)nsu * 1024 * 16 * sizeof(int64_t);
    if (c->gscr_cap < need && c->no_alloc) return rsg_fail(errbuf, "gram: split-K partials need an allocation");
    if (c->gscr_cap < need) {
      if (c->d_gscr) hipFree(c->d_gscr);
      c->d_gscr = nullptr; c->gscr_cap = 0;
      HIPCHK(hipMalloc(&c->d_gscr, need));
      c->gscr_cap = need;
    }
    part = (int64_t *)c->d_gscr;
  }
  const int per = (ndp + 7) / 8;
  int t; rsg_tstart(c, "gram_w8", &t);
  hipLaunchKernelGGL((k_gram_w8<KT>), dim3((unsigned)(8 * per + nsu)), dim3(256), 0, c->stream, m->d_xt, m->d_wq, d_scale,
                     m->Npad, m->Lpad, m->L, out, m->d_nsx, tiles, ndp, per, ntiles, sbase, sextra, part);
  if (nrem > 0)
    hipLaunchKernelGGL(k_gram_w8_fix, dim3((unsigned)(4 * nrem)), dim3(256), 0, c->stream, part, tiles, ndp, sbase, sextra,
                       d_scale, m->L, out, m->d_nsx);
  rsg_tstop(c, t);
  HIPCHK(hipGetLastError());
  return RSG_OK;
}
