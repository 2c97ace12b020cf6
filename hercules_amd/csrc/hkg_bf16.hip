// BF16 MFMA certified lower bound: a batched X.Q^T on the matrix cores
// prunes (query, series) pairs; every survivor is refined exactly in FP32.
//
// For a query q and a series x (FP32 values), with q~, x~ their BF16
// roundings, the MFMAs accumulate, in FP32 from the initial value
// xn2 = |x|^2 (rounded down), the L products x~_i (-2 q~_i) (exact in FP32:
// -2 q~ is a BF16 value and a product of two BF16 values fits FP32), to v:
//   d2(q, x) = |q|^2 + |x|^2 - 2 q.x
//   |q.x - q~.x~| <= |q - q~| |x| + |q~| |x - x~|          (Cauchy-Schwarz)
//   |v - (xn2 - 2 q~.x~)| <= g (xn2 + 2 |q~| |x~|),  g = 2 (L + 1) 2^-24
//       (any summation order of the L + 1 terms; sum |x~_i q~_i| <= |q~||x~|)
// so  LB = |q|^2 + v - (|q - q~| A_x + |q~| B_x + D_x)
// with A_x = 2|x|, B_x = 2 (|x - x~| + g |x~|) and D_x = g xn2 is a lower
// bound of d2 (and the same expression with + is an upper bound).  The norms
// are computed in FP64 and rounded in the safe direction; the FP32
// evaluation is covered by a further absolute slack, and the comparison
// with the k-th distance uses the same certified margin as the other
// bounds.  LB is used only to prune.
//
// The filter tests every pair first against the bound with the largest
// slack terms of its 96-row tile (a minimum per pair: the accumulators
// already hold v); only pairs that pass are re-tested with their own.
#include <float.h>
#include <math.h>
#include <string.h>

#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "hkg_engine.h"
#include "hkg_kernels.h"

namespace hkg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {

// Waves per filter block: one per SIMD, so a wave can hold its query
// operands (8 KS registers) with up to 512 registers including the
// accumulation registers.
__host__ __device__ constexpr int filter_waves(int) { return 4; }
constexpr int kRowTile = 96;               // series per LDS tile
constexpr int kBuf = 3;                    // LDS tile buffers (tiles loaded two ahead)
constexpr uint32_t kRegionCap = 131072;    // 16-byte units per filter block for survivor records (2 MB)
constexpr uint32_t kRecU4 = 5;             // 16-byte units per survivor record
constexpr uint32_t kScatterBlocks = 4;     // k_bf16_scatter blocks per region slice
constexpr float kEvalSlack = 1.0e-3f;      // FP32 evaluation of the bound
constexpr uint32_t kBf16SurvivorCap = 524288;  // survivors kept per query (8 B each)
constexpr float kHistLo = 0.6f;            // histogram range: [0.6 bound, threshold]
constexpr uint32_t kPaceSpins = 4096;      // bounded wait for sibling filter blocks (s_sleep 4 each)

__device__ __forceinline__ uint16_t to_bf16_bits(float f)
{
  uint32_t u = __float_as_uint(f);
  // round to nearest even (finite inputs)
  uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__device__ __forceinline__ float from_bf16_bits(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// rounding bound of the filter's accumulation (L products and |x|^2)
__host__ __device__ constexpr double acc_gamma_d(uint32_t L) { return 2.0 * (double)(L + 1) * 5.9604644775390625e-08; }
__host__ __device__ constexpr float acc_gamma(uint32_t L) { return (float)acc_gamma_d(L); }

__device__ __forceinline__ float rd(double v)
{
  float f = (float)v;
  if ((double)f > v) f = nextafterf(f, -FLT_MAX);
  return f;
}
__device__ __forceinline__ float ru(double v)
{
  float f = (float)v;
  if ((double)f < v) f = nextafterf(f, FLT_MAX);
  return f;
}

// Threshold histogram: bin 0 counts upper bounds below lo, bins 1.. split
// [lo, hi] (hi = the filter threshold) into equal widths.
struct HistRange {
  float lo, w, inv;
};
__device__ __forceinline__ HistRange hist_range(uint32_t bound_bits, float hi)
{
  HistRange r;
  r.lo = kHistLo * __uint_as_float(bound_bits);
  r.w = (hi - r.lo) / (float)(kBf16Bins - 1);
  r.inv = r.w > 0.0f ? 1.0f / r.w : 0.0f;
  return r;
}

// rows -> BF16 copy, {|x|^2 (down), A_x (up), B_x (up), D_x (up)} and |x|^2 alone
__global__ void k_make_bf16(uint64_t n, uint32_t L, const float *__restrict__ X, uint16_t *__restrict__ Xb,
                            float4 *__restrict__ st, float *__restrict__ xn2)
{
  uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const float *x = X + r * L;
  uint16_t *xb = Xb + r * L;
  double n2 = 0.0, e2 = 0.0, b2 = 0.0;
  for (uint32_t t = 0; t < L; t += 2) {
    float a = x[t], c = x[t + 1];
    uint16_t ba = to_bf16_bits(a), bc = to_bf16_bits(c);
    *reinterpret_cast<uint32_t *>(xb + t) = (uint32_t)ba | ((uint32_t)bc << 16);
    double da = (double)from_bf16_bits(ba), dc = (double)from_bf16_bits(bc);
    n2 += (double)a * a + (double)c * c;
    e2 += ((double)a - da) * ((double)a - da) + ((double)c - dc) * ((double)c - dc);
    b2 += da * da + dc * dc;
  }
  const double gL = acc_gamma_d(L);
  const double up = 1.0 + 1e-15;
  const float n2d = rd(n2 * (1.0 - 1e-15));
  st[r] = make_float4(n2d, ru(2.0 * sqrt(n2) * up), ru(2.0 * (sqrt(e2) * up + gL * sqrt(b2) * up) * up),
                      ru(gL * (double)n2d * up));
  xn2[r] = n2d;
}

// per-tile maxima of A_x and B_x
__global__ void k_tile_slack(uint64_t n, const float4 *__restrict__ st, float2 *__restrict__ ts)
{
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t r0 = t * kRowTile;
  if (r0 >= n) return;
  const uint64_t r1 = r0 + kRowTile < n ? r0 + kRowTile : n;
  float a = 0.0f, b = 0.0f;
  for (uint64_t r = r0; r < r1; ++r) {
    a = fmaxf(a, st[r].y);
    b = fmaxf(b, st[r].z);
  }
  ts[t] = make_float2(a, b);
}

// queries -> -2 q~ in BF16 (exact: q~ = bf16(q) scaled by a power of two)
// + {|q|^2 (down), |q - q~| (up), |q~| (up), 0}
__global__ void k_query_bf16(uint32_t nq, uint32_t L, const float *__restrict__ Q, uint16_t *__restrict__ Qb,
                             float4 *__restrict__ st)
{
  uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const float *x = Q + (uint64_t)q * L;
  double n2 = 0.0, e2 = 0.0, b2 = 0.0;
  for (uint32_t t = 0; t < L; ++t) {
    float a = x[t];
    uint16_t b = to_bf16_bits(a);
    Qb[(uint64_t)q * L + t] = to_bf16_bits(-2.0f * from_bf16_bits(b));
    double d = (double)from_bf16_bits(b);
    n2 += (double)a * a;
    e2 += ((double)a - d) * ((double)a - d);
    b2 += d * d;
  }
  st[q] = make_float4(rd(n2 * (1.0 - 1e-15)), ru(sqrt(e2) * (1.0 + 1e-15)), ru(sqrt(b2) * (1.0 + 1e-15)), 0.0f);
}

typedef __attribute__((address_space(3))) void lds_void;

// asynchronous copy of 16 bytes per lane from global memory into LDS at
// lds + 16 * lane (lds is wave-uniform)
__device__ __forceinline__ void load_lds16(const void *g, lds_void *lds)
{
  __builtin_amdgcn_global_load_lds(g, lds, 16, 0, 0);
}
__device__ __forceinline__ void load_lds4(const void *g, lds_void *lds)
{
  __builtin_amdgcn_global_load_lds(g, lds, 4, 0, 0);
}

// Filter.  Block = W waves x 32 QT queries (QT 32-query tiles per wave) of
// block group g (grid.y), wave w taking the queries qperm[32 QT (w G + g) ..]
// (G groups: a group's waves hold slots spread over the batch, so the
// groups' loads even out between pacing points); it loops over the
// row tiles of the batch's tile list (tiles overlapping a leaf whose EAPCA
// bound is within the k-th distance bound of one of the batch's queries).
// Tiles of kRowTile series go through a ring of kBuf LDS buffers, loaded
// two tiles ahead by direct global->LDS copies (the series in BF16, their
// |x|^2 and their slack terms) that each wave issues among its MFMAs.  The
// loop waits on vector memory only for the tile it is about to use: the
// pre-test and the exact bound read LDS, and survivors go to the wave's
// slice of a per-block region in global memory, at positions from ballots
// (k_bf16_scatter files them per query).
//
// MFMA 32x32x16 BF16, D = A.B: A = 32 series x 16 dims from LDS (lane l:
// series l%32, dims 8(l/32)..+8), B = 16 dims x 32 queries from registers
// (lane l: query l%32, same dims); D[i][j] (series i, query j) sits in lane
// j + 32((i/4)%2), register 4(i/8) + i%4.  So every lane owns one query per
// query tile and 16 series per 32-row sub-tile.
//
// LDS tile layout: row r, 16-byte vector v stored at vector position
// v ^ (r & SW) of the row (rows of L bf16, no padding): conflict-free MFMA
// operand reads, and each wave's direct load fills 1 KB of consecutive LDS.
//
// HIST = false: pairs whose certified lower bound is within the query's
// threshold become survivors.  HIST = true: pairs whose certified upper
// bound is within the threshold are counted in the query's histogram
// (global atomics; only pairs passing the lower-bound pre-test get there,
// which undercounts nothing that matters: an undercount only loosens the
// bound).
template <int KS, bool HIST, int QT, int W = filter_waves(KS)>
__global__ void __launch_bounds__(W * 64) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_bf16_filter(uint64_t n, uint32_t nq, const uint16_t *__restrict__ Xb, const float *__restrict__ xn2g,
              const float4 *__restrict__ xst, const float2 *__restrict__ tslack, const uint16_t *__restrict__ Qb,
              const float4 *__restrict__ qst, const uint32_t *__restrict__ bound, const uint32_t *__restrict__ qperm,
              const uint32_t *__restrict__ tiles, uint32_t tile_stride, const uint32_t *__restrict__ ntiles_g,
              uint32_t scap, uint64_t *__restrict__ surv, uint32_t *__restrict__ scnt, const float *__restrict__ thr0,
              uint32_t *__restrict__ hist, uint4 *__restrict__ region, uint32_t region_cap,
              uint32_t *__restrict__ region_cnt, uint32_t *prog, const uint32_t *__restrict__ slot_need,
              uint32_t slot_words, unsigned long long *__restrict__ counters)
{
  constexpr int L = KS * 16;
  constexpr int VPR = L / 8;                       // 16-byte vectors per row
  // rounding of an MFMA chain of L products and the initial |x|^2 (any
  // order): at most kAccGamma (|x|^2 + 2 |q~| |x~|)
  constexpr float kAccGamma = acc_gamma(L);
  // swizzle mask (VPR is a multiple of SW + 1): the 16 lanes that one LDS
  // cycle serves for a 16-byte read are 16 consecutive rows at the same
  // vector position, so 16 distinct positions keep them conflict-free
  constexpr int SW = (VPR % 16 == 0) ? 15 : (VPR % 8 == 0) ? 7 : 3;
  constexpr int NCH = kRowTile * VPR / 64;         // 1 KB chunks per tile
  constexpr int kWaves = W;
  constexpr int QPW = 32 * QT;                     // queries per wave (QT 32-query MFMA tiles)
  // vector memory ops every wave issues per tile, at least: its share of
  // the series chunks and one |x|^2 load (kRowTile / W rows)
  constexpr int kWaitOps = NCH / W + 1;
  static_assert(VPR % (SW + 1) == 0 && (kRowTile * VPR) % 64 == 0 && kRowTile == 96 && W == 4 && kWaitOps <= 15,
                "tile shape");
  // three distinct buffers (not one array): the compiler then sees that the
  // copies in flight into one buffer do not alias reads of another and
  // does not wait for them
  struct TileBuf {
    uint16_t x[kRowTile * L];  // series, BF16, swizzled rows
    float n2[kRowTile];        // |x|^2 (rounded down), the accumulators' initial values
  };
  __shared__ __attribute__((aligned(16))) TileBuf tb0, tb1, tb2;
  auto buf = [&](auto Bc) -> TileBuf & {
    constexpr int B = decltype(Bc)::value;
    if constexpr (B == 0)
      return tb0;
    else if constexpr (B == 1)
      return tb1;
    else
      return tb2;
  };
  const int lane = threadIdx.x & 63, half = lane >> 5, l32 = lane & 31;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar loops and copy bases
  const uint32_t g = blockIdx.y;
  const uint32_t blk = blockIdx.y * gridDim.x + blockIdx.x;
  // this wave's slot: QPW consecutive queries of the batch order (waves are
  // dealt to the block groups round-robin)
  const uint32_t wslot = wave * gridDim.y + g;

  // this lane's query of each query tile, and its constants
  uint32_t qid[QT];
  float qn2[QT], qe[QT], qbn[QT], thr[QT], hlo[QT], hinv[QT];
  bf16x8 qop[QT][KS];
#pragma unroll
  for (int t = 0; t < QT; ++t) {
    // the batch is ordered by best leaf, so a wave holds neighbouring
    // queries (their survivors fall in the same tiles); waves are dealt to
    // the block groups round-robin, so each group gets a share of every
    // region and the groups' survivor loads even out
    const uint32_t slot = wslot * QPW + t * 32 + l32;
    const uint32_t q = slot < nq ? qperm[slot] : 0xFFFFFFFFu;
    qid[t] = q;
    const uint32_t qs = q < nq ? q : 0;
    const float4 s = qst[qs];
    qn2[t] = s.x;
    qe[t] = s.y;
    qbn[t] = s.z;
    thr[t] = q < nq ? thr0[q] : -FLT_MAX;  // padding never passes
    hlo[t] = 0.0f;
    hinv[t] = 0.0f;
    if (HIST && q < nq) {
      const uint32_t bb = bound[q];
      if (bb >= kInfBits) {
        thr[t] = -FLT_MAX;  // no finite bound: nothing to histogram
      } else {
        const HistRange r = hist_range(bb, thr[t]);
        hlo[t] = r.lo;
        hinv[t] = r.inv;
      }
    }
    const bf16x8 *src = reinterpret_cast<const bf16x8 *>(Qb + (uint64_t)qs * L);
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      qop[t][k] = src[2 * k + half];
      // the query operands live in accumulation registers (the MFMAs read
      // their B operand from there); the accumulators, built with
      // -amdgpu-mfma-vgpr-form, stay in vector registers where the
      // pre-test reads them without copies
      asm volatile("" : "+a"(qop[t][k]));
    }
  }
  // survivor records of this wave: its slice of the block's region
  // (region_cap / W 16-byte units, kRecU4 per record), counted in a
  // wave-uniform register (positions from ballots, no atomics)
  const uint32_t wcap = region_cap / kWaves;
  const uint32_t wrec = wcap / kRecU4;
  const uint64_t wbase = ((uint64_t)blk * kWaves + wave) * wcap;
  uint32_t wcnt = 0;

  // every block group walks the one tile list of the batch (tile_stride =
  // 0) or its own; block x takes list entries x, x + grid.x, ..., so block
  // x of every group reads the same tiles at about the same time on the
  // same XCD
  const uint32_t gl = tile_stride ? g : 0;
  const uint32_t nt = ntiles_g[gl];
  const uint32_t my_n = blockIdx.x < nt ? (nt - blockIdx.x + gridDim.x - 1) / gridDim.x : 0;
  const uint32_t *tlist = tiles + (uint64_t)gl * tile_stride;
  // the block's tile ids and slack maxima in registers, 64 tiles per
  // chunk, one per lane: the chunk of the current step (cid, cts) and the
  // next (nid, nts), read with readlane.  No memory access in the steps:
  // scalar loads would share the LDS wait counter and return out of
  // order, and an LDS read after the tile copies are issued makes the
  // compiler wait for all of them.  A chunk is fetched and waited for at
  // once (the empty asm uses its registers), so that no load is pending
  // across steps: the compiler would otherwise drain every copy in flight
  // before touching those registers.
  auto id_of = [&](uint32_t c) -> uint32_t {
    const uint32_t j = c * 64 + lane;
    return j < my_n ? tlist[blockIdx.x + (uint64_t)j * gridDim.x] : 0u;
  };
  // slot_need (nullable): per tile, one bit per wave slot, set when a query
  // of the slot needs a leaf the tile overlaps (EAPCA bound within its k-th
  // distance bound); a wave whose bit is clear skips the tile's MFMAs.  The
  // wave's bit rides in bit 31 of the tile id it holds (tile ids are far
  // below 2^31), so it costs no register.
  const uint32_t nw_off = wslot >> 5, nw_bit = wslot & 31;
  auto fetch_chunk = [&](uint32_t c, uint32_t &id, float2 &ts) {
    id = id_of(c);
    ts = tslack[id];
    const uint32_t nw = slot_need ? slot_need[(uint64_t)id * slot_words + nw_off] : 0xFFFFFFFFu;
    id |= ((nw >> nw_bit) & 1u) << 31;
    asm volatile("" ::"v"(id), "v"(ts.x), "v"(ts.y));
  };
  uint32_t cid, nid;
  float2 cts, nts;
  fetch_chunk(0, cid, cts);
  fetch_chunk(1, nid, nts);
  // id of the block's j-th tile, from the chunk of step jc or the next
  auto tile_at = [&](uint32_t j, uint32_t jc) -> uint32_t {
    return ((j >> 6) == (jc >> 6) ? __builtin_amdgcn_readlane(cid, j & 63) : __builtin_amdgcn_readlane(nid, j & 63)) &
           0x7FFFFFFFu;
  };
  auto slack_at = [&](uint32_t j) -> float2 {
    return make_float2(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(cts.x), j & 63)),
                       __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cts.y), j & 63)));
  };
  // Copies of the tile at rows r0.. into buffer Bc: 1 KB series chunks c =
  // wave + i W (this wave's i-th), and one load of the |x|^2 of kRowTile / W
  // rows per wave
  constexpr int NCI = (NCH + kWaves - 1) / kWaves;  // series chunks per wave, at most
  constexpr int kN2Rows = kRowTile / kWaves;       // |x|^2 rows per wave
  auto issue_chunk = [&](uint64_t r0, auto Bc, int i) __attribute__((always_inline)) {
    TileBuf &tbuf = buf(Bc);
    const int c = wave + i * kWaves;
    if (NCH % kWaves == 0 || c < NCH) {
      const int lin = c * 64 + lane;  // LDS vector index of this lane
      const int r = lin / VPR, p = lin - r * VPR;
      // a uniform base and a 32-bit lane offset (the rows are padded to
      // whole tiles, so the last tile needs no clamp)
      const uint16_t *src = Xb + r0 * L + (uint32_t)(r * L + (p ^ (r & SW)) * 8);
      load_lds16(src, (lds_void *)(reinterpret_cast<char *>(&tbuf.x[0]) + c * 1024));
    }
  };
  auto issue_stats = [&](uint64_t r0, auto Bc) __attribute__((always_inline)) {
    TileBuf &tbuf = buf(Bc);
    const int rbase = wave * kN2Rows;
    if (lane < kN2Rows)
      load_lds4(xn2g + r0 + rbase + lane, (lds_void *)(reinterpret_cast<char *>(&tbuf.n2[0]) + rbase * 4));
  };
  auto issue = [&](uint32_t tile, auto Bc) __attribute__((always_inline)) {
    const uint64_t r0 = (uint64_t)tile * kRowTile;
#pragma unroll
    for (int i = 0; i < NCI; ++i) issue_chunk(r0, Bc, i);
    issue_stats(r0, Bc);
  };

  // The accumulators of a sub-tile start at the |x|^2 of their series
  // (register r: series crow + 8(r/4) + 4 half + r%4 of the tile) and the
  // query operands hold -2 q~, so an MFMA chain ends at v = xn2 - 2 s: the
  // pre-test needs no arithmetic but a minimum.  m_t = min over the lane's 16
  // series of v for query tile t; some pair passes (v <= C_t) iff m_t <= C_t,
  // as fminf skips NaN the way <= rejects it.
  auto n2_of = [&](auto Bc, int crow, f32x16 &c) __attribute__((always_inline)) {
    const TileBuf &tbuf = buf(Bc);
#pragma unroll
    for (int gi = 0; gi < 4; ++gi) {
      const float4 x4 = *reinterpret_cast<const float4 *>(&tbuf.n2[crow + 8 * gi + 4 * half]);
      c[4 * gi] = x4.x;
      c[4 * gi + 1] = x4.y;
      c[4 * gi + 2] = x4.z;
      c[4 * gi + 3] = x4.w;
    }
  };
  using Acc = f32x16[QT];
  auto pt_elem = [&](int r, const Acc &p, float (&m)[QT]) __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < QT; ++t) m[t] = r == 0 ? p[t][r] : fminf(m[t], p[t][r]);
  };
  // MFMAs of 32-row sub-tile SUB (rows crow = 32 SUB.. of buffer Bc) into
  // a[0..QT), all starting from cin (the sub-tile's |x|^2).  The series
  // operands are read in batches of G k-steps, each batch issued right after
  // the first MFMA of the batch before: with the tile copies in flight the
  // compiler can only wait for all LDS reads at once, so every read gets most
  // of a batch of MFMAs to land.  The first batch of the next sub-tile of the
  // same tile and its |x|^2 are read the same way during the last batch
  // (NEXT), into opc and cin, so that they are there when that sub-tile starts
  // (PRE); only the first sub-tile of a tile waits for its first reads.
  // Between the MFMAs this wave issues its copies of the tile at rows r0n..
  // into buffer BNc, a third per sub-tile, at batch starts: all waves issuing
  // a whole tile at once would queue on the CU's copy path and stall, while
  // spread over the MFMAs the copies proceed in the background.  With PT,
  // the pre-test of the previous sub-tile (accumulators p) runs between
  // the MFMAs, from the third k-step on (its accumulators are complete by
  // then): the matrix cores stay busy while the wave, alone on its SIMD, does
  // the vector work.
  constexpr int G = KS % 4 == 0 ? 4 : (KS % 3 == 0 ? 3 : 2);  // k-steps per operand batch
  constexpr int NB = KS / G;
  static_assert(KS % G == 0, "operand batches");
  auto mfma_sub = [&](auto Bc, auto SUBc, Acc &a, auto PTc, const Acc &p, float (&m)[QT], uint64_t r0n, auto BNc,
                      bf16x8 (&opc)[G], f32x16 &cin, auto PREc, auto NEXTc) __attribute__((always_inline)) {
    constexpr bool PT = decltype(PTc)::value;
    constexpr int SUB = decltype(SUBc)::value;
    constexpr bool PRE = decltype(PREc)::value, NEXT = decltype(NEXTc)::value;
    constexpr int c_lo = SUB * NCI / 3, c_n = (SUB + 1) * NCI / 3 - c_lo;  // this sub-tile's copy chunks
    const int crow = 32 * SUB;
    const bf16x8 *xv = reinterpret_cast<const bf16x8 *>(&buf(Bc).x[0]);
    const int row = crow + l32, nrow = row + 32;
    auto op = [&](int k) { return xv[row * VPR + ((2 * k + half) ^ (row & SW))]; };
    auto op_next = [&](int k) { return xv[nrow * VPR + ((2 * k + half) ^ (nrow & SW))]; };
    // pre-test elements done after k-step k: 0 for k < 2, then spread evenly
    auto done = [](int k) { return k < 2 ? 0 : (k - 2) * 16 / (KS - 2); };
    __builtin_amdgcn_sched_barrier(0);
    bf16x8 opn[G];
    if constexpr (!PRE) {
#pragma unroll
      for (int g = 0; g < G; ++g) opc[g] = op(g);
      n2_of(Bc, crow, cin);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int k = b * G + g;
#pragma unroll
        for (int t = 0; t < QT; ++t)
          a[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(opc[g], qop[t][k], k == 0 ? cin : a[t], 0, 0, 0);
        if (g == 0) {
          // copies: chunk c_lo + q at batch q NB / c_n; |x|^2 at batch 0 of
          // sub-tile 0
#pragma unroll
          for (int q = 0; q < c_n; ++q)
            if (b == q * NB / c_n) issue_chunk(r0n, BNc, c_lo + q);
          if (SUB == 0 && b == 0) issue_stats(r0n, BNc);
          if (b + 1 < NB) {
#pragma unroll
            for (int h = 0; h < G; ++h) opn[h] = op((b + 1) * G + h);
          } else if constexpr (NEXT) {
#pragma unroll
            for (int h = 0; h < G; ++h) opn[h] = op_next(h);
            n2_of(Bc, crow + 32, cin);
          }
        }
        if constexpr (PT) {
#pragma unroll
          for (int r = done(k); r < done(k + 1); ++r) pt_elem(r, p, m);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      if (b + 1 < NB || NEXT) {
#pragma unroll
        for (int g = 0; g < G; ++g) opc[g] = opn[g];
      }
    }
  };
  auto pretest = [&](const Acc &a, float (&m)[QT]) __attribute__((always_inline)) {
#pragma unroll
    for (int r = 0; r < 16; ++r) pt_elem(r, a, m);
  };
  // Pairs of a sub-tile whose pre-test passed (rare; reads no LDS, so it
  // may run after the tile's buffer is reused); only the query tiles t whose
  // ballot hm[t] (lanes with m_t <= C_t) is not empty are looked at.  HIST:
  // the pairs whose upper bound with their own slack is within the threshold
  // go to the query's histogram, one ballot per pair.  Otherwise each lane of
  // the ballot writes one record to the wave's slice of the region: {series
  // of its register 0, query, slot, 0} and its 16 values xn2 - 2 s, five
  // 16-byte stores at a position from the ballot; k_bf16_scatter tests each
  // of the 16 pairs with its own slack.  A wave that finds a survivor spends
  // a few instructions and stores on it, so it holds up its block's barrier
  // only briefly.
  auto slow = [&](int crow, uint64_t r0, const Acc &a, const float (&C)[QT], const uint64_t (&hm)[QT]) __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < QT; ++t) {
      if (hm[t] == 0) continue;
      const float Ct = C[t];
      if constexpr (HIST) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const float v = a[t][r];  // xn2 - 2 s
          const bool pass = v <= Ct;
          if (__builtin_expect(__ballot(pass) == 0, 1)) continue;
          const uint64_t grow = r0 + crow + 8 * (r >> 2) + 4 * half + (r & 3);
          if (pass && grow < n) {
            const float4 xsv = xst[grow];
            const float ub = (v + qn2[t]) + (qe[t] * xsv.y + qbn[t] * xsv.z + xsv.w + kEvalSlack);
            if (ub <= thr[t]) {
              const float lo = hlo[t];
              const int bin = ub < lo ? 0 : min(kBf16Bins - 1, 1 + (int)fminf((ub - lo) * hinv[t], (float)kBf16Bins));
              atomicAdd(&hist[(uint64_t)qid[t] * kBf16Bins + bin], 1u);
            }
          }
        }
      } else {
        const uint64_t m = hm[t];
        const uint32_t pos = wcnt + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        wcnt += (uint32_t)__popcll(m);
        const uint64_t rowb = r0 + crow + 4 * half;  // series of register 0; register r: + 8 (r / 4) + r % 4
        if ((m >> lane) & 1u) {
          if (pos < wrec) {
            uint4 *rec = region + wbase + (uint64_t)pos * kRecU4;
            rec[0] = make_uint4((uint32_t)rowb, qid[t], (uint32_t)(t * 32 + l32), 0u);
#pragma unroll
            for (int gi = 0; gi < 4; ++gi)
              rec[1 + gi] = make_uint4(__float_as_uint(a[t][4 * gi]), __float_as_uint(a[t][4 * gi + 1]),
                                       __float_as_uint(a[t][4 * gi + 2]), __float_as_uint(a[t][4 * gi + 3]));
          } else {
            // the wave's slice is full: tested here, straight to the query's list
#pragma unroll
            for (int r = 0; r < 16; ++r) {
              const uint64_t grow = rowb + 8 * (r >> 2) + (r & 3);
              if (a[t][r] <= Ct && grow < n) {
                const float4 xsv = xst[grow];
                const float lb = (a[t][r] + qn2[t]) - (qe[t] * xsv.y + qbn[t] * xsv.z + xsv.w + kEvalSlack);
                if (lb <= thr[t]) {
                  const uint32_t qp = atomicAdd(&scnt[qid[t]], 1u);
                  if (qp < scap)
                    surv[(uint64_t)qid[t] * scap + qp] = ((uint64_t)__float_as_uint(fmaxf(lb, 0.0f)) << 32) | (uint32_t)grow;
                }
              }
            }
          }
        }
      }
    }
  };
  using B0 = std::integral_constant<int, 0>;
  using B1 = std::integral_constant<int, 1>;
  using B2 = std::integral_constant<int, 2>;
  if (my_n > 0) issue(tile_at(0, 0), B0{});
  if (my_n > 1) issue(tile_at(1, 0), B1{});
  // one tile: j in buffer Bc (= j % 3); issues tile j + 2 into the buffer
  // that held tile j - 1
  // Sibling blocks (the same x in the other block groups, so on the same
  // XCD) walk the same tiles.  Every 128 tiles a block publishes its
  // progress and waits until its siblings have reached the same tile, so
  // that one block's copies from HBM are the others' L2 hits and the
  // batch reads the series about once instead of once per group.  The
  // wait is bounded: a sibling that is not resident (CUs held by another
  // kernel) costs one timeout, after which the block stops waiting.
  // Results never depend on it.
  const bool pace_grid = tile_stride == 0 && gridDim.y > 1;
  bool pace = pace_grid;  // read by wave 0 only
  auto keep_pace = [&](uint32_t j) __attribute__((always_inline)) {
    if (!pace_grid) return;  // uniform over the block
    if (wave == 0 && pace) {
      if (lane == 0) __hip_atomic_store(&prog[blk], j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const bool sib = (uint32_t)lane < gridDim.y && (uint32_t)lane != blockIdx.y;
      const uint32_t *p = prog + (sib ? lane * gridDim.x + blockIdx.x : 0u);
      uint32_t spins = 0;
      bool behind;
      do {
        const uint32_t v = sib ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0xFFFFFFFFu;
        behind = __ballot(v < j) != 0;
        if (behind) __builtin_amdgcn_s_sleep(4);
      } while (behind && ++spins < kPaceSpins);
      if (behind) pace = false;
    }
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  };
  // sub-tile 2 of the previous step (accumulators, pre-test constants,
  // first row), tested during the next step's first MFMAs; nothing passes
  // before the first tile
  Acc Y;
  float Cp[QT];
#pragma unroll
  for (int t = 0; t < QT; ++t) {
    Y[t] = f32x16{};
    Cp[t] = -FLT_MAX;
  }
  uint64_t r0p = 0;
  // the lanes of each query tile with a pair within their pre-test constant
  auto ballots = [&](const float (&m)[QT], const float (&C)[QT], uint64_t (&hm)[QT]) __attribute__((always_inline)) {
    uint64_t any = 0;
#pragma unroll
    for (int t = 0; t < QT; ++t) {
      hm[t] = __ballot(m[t] <= C[t]);
      any |= hm[t];
    }
    return any != 0;
  };
  uint32_t skipped = 0;  // steps this wave skipped
  auto step = [&](auto Bc, uint32_t j) __attribute__((always_inline)) {
    constexpr int B = decltype(Bc)::value;
    using BN = std::integral_constant<int, (B + 2) % kBuf>;
    // this wave's loads of tile j have landed (those of tile j + 1 may still
    // be in flight), then every wave's; tile j - 1 is no longer read
    if (j + 1 < my_n)
      __builtin_amdgcn_s_waitcnt(0xF70 | kWaitOps);
    else
      __builtin_amdgcn_s_waitcnt(0);
    // a plain barrier (and a compiler memory barrier): the fences of
    // __syncthreads would also wait for tile j + 1's copies, as those
    // write LDS.  Loads complete in order, so the wait above covers every
    // copy of tile j even with survivor stores in flight.
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    // a new chunk: the next one becomes current and the one after is
    // fetched (only tile j + 1's copies are in flight here)
    if ((j & 63) == 0 && j > 0) {
      cid = nid;
      cts = nts;
      fetch_chunk((j >> 6) + 1, nid, nts);
      if ((j & 127) == 0) keep_pace(j);
    }
    // the tile two ahead, copied during this step's MFMAs (past the end
    // of the list, tile 0 again into the free buffer: the same count of
    // copies every step, and no branch among the MFMAs)
    const uint64_t r0n = (uint64_t)tile_at(j + 2, j) * kRowTile;
    const uint64_t r0 = (uint64_t)tile_at(j, j) * kRowTile;
    if (!(__builtin_amdgcn_readlane(cid, j & 63) >> 31)) {
      // no query of this wave needs the tile: its copies of tile j + 2 (the
      // same count as a computing step issues), the previous step's sub-tile
      // 2, and no MFMAs
#pragma unroll
      for (int i = 0; i < NCI; ++i) issue_chunk(r0n, BN{}, i);
      issue_stats(r0n, BN{});
      float m[QT];
      uint64_t h[QT];
      pretest(Y, m);
      if (__builtin_expect(ballots(m, Cp, h), 0)) slow(64, r0p, Y, Cp, h);
#pragma unroll
      for (int t = 0; t < QT; ++t) Cp[t] = -FLT_MAX;  // nothing pending for the next step
      ++skipped;
      return;
    }
    using PT = std::true_type;
    using T_ = std::true_type;
    using F_ = std::false_type;
    using S0 = std::integral_constant<int, 0>;
    using S1 = std::integral_constant<int, 1>;
    using S2 = std::integral_constant<int, 2>;
    Acc A, X;
    float m[QT], C[QT];
    uint64_t h[QT];
    // sub-tiles 0..2, the MFMAs of one beside the pre-test of the one
    // before (for sub-tile 0: the previous step's sub-tile 2)
    bf16x8 opc[G];  // first operand batch of the next sub-tile
    f32x16 cin;     // |x|^2 of the next sub-tile
    mfma_sub(Bc, S0{}, A, PT{}, Y, m, r0n, BN{}, opc, cin, F_{}, T_{});
    if (__builtin_expect(ballots(m, Cp, h), 0)) slow(64, r0p, Y, Cp, h);
    // pre-test constant: xn2 - 2 s <= C  <=>  bound with the tile's largest
    // slack terms <= threshold (the accumulation term g |x|^2 at most
    // g (A_x / 2)^2)
    const float2 tsl = slack_at(j);
    const float dmax = kAccGamma * (0.25f * tsl.x * tsl.x);
#pragma unroll
    for (int t = 0; t < QT; ++t) C[t] = (thr[t] - qn2[t]) + (qe[t] * tsl.x + qbn[t] * tsl.y + dmax + kEvalSlack);
    mfma_sub(Bc, S1{}, X, PT{}, A, m, r0n, BN{}, opc, cin, T_{}, T_{});
    if (__builtin_expect(ballots(m, C, h), 0)) slow(0, r0, A, C, h);
    mfma_sub(Bc, S2{}, Y, PT{}, X, m, r0n, BN{}, opc, cin, T_{}, F_{});
    if (__builtin_expect(ballots(m, C, h), 0)) slow(32, r0, X, C, h);
    // sub-tile 2 is tested during the next step (from its accumulators, so
    // after the next barrier frees this buffer)
#pragma unroll
    for (int t = 0; t < QT; ++t) Cp[t] = C[t];
    r0p = r0;
  };
  static_assert(kBuf == 3, "the loop below cycles three buffers");
  for (uint32_t j = 0; j < my_n; j += 3) {
    step(B0{}, j);
    if (j + 1 < my_n) step(B1{}, j + 1);
    if (j + 2 < my_n) step(B2{}, j + 2);
  }
  // the last tile's sub-tile 2
  {
    float m[QT];
    uint64_t h[QT];
    pretest(Y, m);
    if (ballots(m, Cp, h)) slow(64, r0p, Y, Cp, h);
  }
  // the last steps' copies (re-reads into free buffers) land before the
  // block's LDS is released; siblings no longer wait for this block
  __builtin_amdgcn_s_waitcnt(0);
  if (pace_grid && wave == 0 && lane == 0)
    __hip_atomic_store(&prog[blk], 0xFFFFFFFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (!HIST && lane == 0) region_cnt[(uint64_t)blk * kWaves + wave] = wcnt < wrec ? wcnt : wrec;
  // (query, series) pairs of this wave's real queries it skipped
  const uint32_t nreal = nq > wslot * QPW ? min((uint32_t)QPW, nq - wslot * QPW) : 0u;
  if (counters && lane == 0 && skipped && nreal)
    atomicAdd(&counters[kCtrBf16Skip], (unsigned long long)skipped * nreal * kRowTile);
}

// The survivor records every filter wave wrote (one region slice each, one
// record per lane and query tile whose pre-test passed, with the 16 values
// xn2 - 2 s of the lane's series): each pair's lower bound with the row's own
// slack, and the pairs within the query's threshold into its list, keyed by
// that bound.  One thread per record (usually one or two of its 16 pairs
// pass); a global atomic per kept pair.  A value above the filter's own
// pre-test constant for the row tile (the tile's largest slack terms, the
// same float arithmetic as the filter) cannot pass with the row's own
// slack, so its row's terms are not read.
__global__ void __launch_bounds__(256) k_bf16_scatter(const uint4 *__restrict__ region, uint32_t slice_cap,
                                                      const uint32_t *__restrict__ region_cnt, uint64_t n,
                                                      const float4 *__restrict__ xst, const float2 *__restrict__ tslack,
                                                      float acc_g, const float4 *__restrict__ qst,
                                                      const float *__restrict__ thr0, uint32_t scap,
                                                      uint64_t *__restrict__ surv, uint32_t *__restrict__ scnt)
{
  const uint32_t sl = blockIdx.x;
  const uint32_t cnt = min(region_cnt[sl], slice_cap / kRecU4);
  const uint4 *reg = region + (uint64_t)sl * slice_cap;
  // grid.y blocks share a slice (the slices' record counts are uneven)
  for (uint32_t i = blockIdx.y * blockDim.x + threadIdx.x; i < cnt; i += blockDim.x * gridDim.y) {
    const uint4 *rec = reg + (uint64_t)i * kRecU4;
    const uint4 h = rec[0];
    const uint32_t q = h.y;
    const float4 qs = qst[q];
    const float th = thr0[q];
    const float2 tsl = tslack[h.x / kRowTile];
    const float dmax = acc_g * (0.25f * tsl.x * tsl.x);
    const float C = (th - qs.x) + (qs.y * tsl.x + qs.z * tsl.y + dmax + kEvalSlack);
#pragma unroll
    for (int gi = 0; gi < 4; ++gi) {
      const uint4 vv = rec[1 + gi];
      const uint32_t vb[4] = {vv.x, vv.y, vv.z, vv.w};
#pragma unroll
      for (int i4 = 0; i4 < 4; ++i4) {
        const uint64_t row = (uint64_t)h.x + 8 * gi + i4;
        if (row >= n) continue;  // a padding row of the last tile
        const float v = __uint_as_float(vb[i4]);
        if (!(v <= C)) continue;
        const float4 xsv = xst[row];
        const float lb = (v + qs.x) - (qs.y * xsv.y + qs.z * xsv.z + xsv.w + kEvalSlack);
        if (lb <= th) {
          const uint32_t p = atomicAdd(&scnt[q], 1u);
          if (p < scap) surv[(uint64_t)q * scap + p] = ((uint64_t)__float_as_uint(fmaxf(lb, 0.0f)) << 32) | (uint32_t)row;
        }
      }
    }
  }
}

// New k-th distance bound of every query from its histogram: the first bin
// where the count reaches k.  Every series counted up to bin j has an upper
// bound below lo + (j + 1) w (one bin of margin for the rounding of the bin
// index), so at least k series lie within that distance.  The bound is
// counted over series, not distinct distances: the refine stage certifies
// the answer against it and flags the query otherwise.
__global__ void k_bf16_tau(uint32_t nq, const uint32_t *__restrict__ kq, const uint32_t *__restrict__ hist,
                           const float *__restrict__ thr0, uint32_t *__restrict__ bound)
{
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const uint32_t bb = bound[q];
  if (bb >= kInfBits) return;
  const HistRange r = hist_range(bb, thr0[q]);
  const uint32_t k = kq[q];
  uint32_t cum = 0;
  for (int j = 0; j < kBf16Bins; ++j) {
    cum += hist[(uint64_t)q * kBf16Bins + j];
    if (cum >= k) {
      const float t = ru((double)r.lo + (double)(j + 1) * (double)r.w);
      if (t < __uint_as_float(bb)) bound[q] = __float_as_uint(t);
      return;
    }
  }
}

}  // namespace

// 32-query MFMA tiles per filter wave: the launch shapes compiled for each
// length (the query operands, KS registers per tile, share the wave's 256
// accumulation registers with nothing else; at L = 256 the accumulators of
// four tiles would not fit beside them)
bool bf16_query_tiles_ok(uint32_t L, uint32_t qt) { return qt == 2 || qt == 3 || (qt == 4 && L <= 128); }

// queries per filter block (block group) for series of length L
static uint32_t bf16_qpb(uint32_t L, uint32_t qt) { return (uint32_t)filter_waves((int)L / 16) * 32 * qt; }
uint32_t bf16_group_queries(uint32_t L, uint32_t qt) { return bf16_qpb(L, qt); }

uint64_t bf16_scratch_per_query(uint32_t L, uint64_t nleaf, uint64_t ntiles)
{
  // ... + the block groups' leaf needs and tile lists, the 32-query slots'
  // leaf needs and the wave slots' tile bits
  return (uint64_t)kBf16SurvivorCap * 8 + 4 + (uint64_t)L * 2 + 16 + 16 + 4 + 4ull * kBf16Bins +
         (nleaf + 4 * ntiles) / bf16_qpb(L, kBf16QueryTiles) + nleaf / 32 + ntiles / 16 + 18;
}

// one resident block per CU, split over the block groups; a multiple of 8
// (the XCD count) so that block x of every group lands on the same XCD
// (blocks go to XCDs round-robin by linear id x + grid.x * y)
static uint32_t bf16_grid_x(uint32_t groups)
{
  int d = 0, cus = 0;
  (void)hipGetDevice(&d);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d);
  const uint32_t per = (uint32_t)std::max(cus, 1) / std::max<uint32_t>(groups, 1);
  return per >= 8 ? per / 8 * 8 : std::max<uint32_t>(per, 1);
}

// shared_list: every block group walks list 0 of tiles / ntiles_g
void launch_bf16_filter(hipStream_t st, const DevIndex &dv, uint32_t nq, const uint16_t *Qb, const float4 *qst,
                        const uint32_t *bound, const uint32_t *qperm, const uint32_t *tiles,
                        const uint32_t *ntiles_g, bool shared_list, uint32_t scap, uint64_t *surv, uint32_t *scnt,
                        const float *thr0, uint32_t *hist, uint4 *region, uint32_t *region_cnt,
                        uint32_t *prog, uint32_t qt, const uint32_t *slot_need, uint32_t slot_words,
                        unsigned long long *counters)
{
  if (hist || !bf16_query_tiles_ok(dv.L, qt)) qt = 2;
  const uint32_t stride = shared_list ? 0 : dv.ntiles;
  const uint32_t qpb = bf16_qpb(dv.L, qt);
  const uint32_t groups = (nq + qpb - 1) / qpb;
  dim3 grid(bf16_grid_x(groups), groups);
  launch_fill(st, prog, grid.x * grid.y, 0);
#define HKG_BF16_LAUNCH(KS, HIST, QT)                                                                             \
  hipLaunchKernelGGL((k_bf16_filter<KS, HIST, QT>), grid, dim3(filter_waves(KS) * 64), 0, st, dv.n, nq, dv.Xb,   \
                     dv.xn2, dv.xst, dv.tile_slack, Qb, qst, bound, qperm, tiles, stride, ntiles_g, scap, surv,   \
                     scnt, thr0, hist, region, kRegionCap, region_cnt, prog, slot_need, slot_words, counters)
  // the histogram pass (rare: the bound pass) runs in the default shape only
#define HKG_BF16_LAUNCH_QT(KS)                   \
  if (hist)                                      \
    HKG_BF16_LAUNCH(KS, true, 2);                \
  else if (qt == 3)                              \
    HKG_BF16_LAUNCH(KS, false, 3);               \
  else if (qt == 4 && KS <= 8)                   \
    HKG_BF16_LAUNCH(KS, false, (KS <= 8 ? 4 : 2)); \
  else                                           \
    HKG_BF16_LAUNCH(KS, false, 2)
  switch (dv.L) {
  case 96: HKG_BF16_LAUNCH_QT(6); break;
  case 128: HKG_BF16_LAUNCH_QT(8); break;
  default: HKG_BF16_LAUNCH_QT(16); break;
  }
#undef HKG_BF16_LAUNCH_QT
#undef HKG_BF16_LAUNCH
  if (!hist) {
    const uint32_t w = (uint32_t)filter_waves((int)dv.L / 16);
    hipLaunchKernelGGL(k_bf16_scatter, dim3(grid.x * grid.y * w, kScatterBlocks), dim3(256), 0, st, region, kRegionCap / w,
                       region_cnt, dv.n, dv.xst, dv.tile_slack, acc_gamma(dv.L), qst, thr0, scap, surv, scnt);
  }
}

// filter blocks of a launch at most (one per CU)
static uint32_t bf16_max_blocks()
{
  int d = 0, cus = 0;
  (void)hipGetDevice(&d);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d);
  return (uint32_t)std::max(cus, 1);
}

// filter threshold of every query: its k-th distance bound with the
// certified margin
__global__ void k_bf16_thr_init(uint32_t nq, const uint32_t *__restrict__ bound, float *__restrict__ thr0)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nq) {
    const uint32_t bb = bound[i];
    thr0[i] = bb >= kInfBits ? FLT_MAX : __uint_as_float(bb) * (1.0f + kLbRel) + kLbAbs;
  }
}

// leaves some query of a block group still needs: EAPCA bound within the
// query's current k-th distance bound
// pairs (nullable): also count every (query, leaf) pair of the group whose
// bound is within the query's k-th bound (then no early exit per leaf)
__global__ void k_group_leaf_need(uint32_t nq, uint32_t qpb, uint32_t nleaf, const float *__restrict__ lb,
                                  const uint32_t *__restrict__ bound, const uint32_t *__restrict__ qperm,
                                  uint8_t *__restrict__ need, unsigned long long *__restrict__ pairs)
{
  const uint32_t leaf = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t g = blockIdx.y;
  uint32_t cnt = 0;
  if (leaf < nleaf) {
    uint8_t nd = 0;
    const uint32_t s0 = g * qpb;
    const uint32_t s1 = min(nq, s0 + qpb);
    for (uint32_t sl = s0; sl < s1 && (pairs || !nd); ++sl) {
      const uint32_t q = qperm[sl];
      const uint32_t bb = bound[q];
      const float v = lb[(uint64_t)q * nleaf + leaf];
      const bool in = bb >= kInfBits || v <= __uint_as_float(bb) * (1.0f + kLbRel) + kLbAbs;
      nd |= in ? 1 : 0;
      cnt += in ? 1u : 0u;
    }
    need[(uint64_t)g * nleaf + leaf] = nd;
  }
  if (pairs) {
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(pairs, (unsigned long long)cnt);
  }
}

// row tiles of each block group: tiles overlapping a needed leaf
__global__ void k_group_tiles(uint32_t ntiles, uint32_t nleaf, const uint32_t *__restrict__ tile_leaf0,
                              const uint32_t *__restrict__ tile_leaf1, const uint8_t *__restrict__ need,
                              uint32_t *__restrict__ tiles, uint32_t *__restrict__ ntiles_g)
{
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t g = blockIdx.y;
  bool keep = false;
  if (t < ntiles) {
    for (uint32_t lf = tile_leaf0[t]; lf <= tile_leaf1[t] && !keep; ++lf) keep = need[(uint64_t)g * nleaf + lf] != 0;
  }
  const uint64_t m = __ballot(keep);
  const int lane = threadIdx.x & 63;
  if (m) {
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&ntiles_g[g], (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (keep) tiles[(uint64_t)g * ntiles + base + __popcll(m & ((1ull << lane) - 1))] = t;
  }
}

// batch order by best leaf (leaves are in tree order, so queries of one
// block group need overlapping sets of leaves): keys = best leaf, values =
// batch index
__global__ void k_iota_best(uint32_t nq, const uint32_t *__restrict__ best, uint32_t *__restrict__ keys,
                            uint32_t *__restrict__ vals)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nq) {
    keys[i] = best[(uint64_t)i * kBestLeaves];
    vals[i] = i;
  }
}

hkg_status Index::make_bf16()
{
  std::lock_guard<std::mutex> lock(mu_bf16);
  if (bf16_ready) return HKG_OK;
  // an earlier attempt that failed part-way
  for (void *p : {(void *)dev.Xb, (void *)dev.xst, (void *)dev.xn2, (void *)dev.tile_slack, (void *)dev.tile_leaf0,
                  (void *)dev.tile_leaf1})
    if (p) (void)hipFree(p);
  dev.Xb = nullptr;
  dev.xst = nullptr;
  dev.xn2 = nullptr;
  dev.tile_slack = nullptr;
  dev.tile_leaf0 = dev.tile_leaf1 = nullptr;
  if (dev.L % 16 != 0 || (dev.L != 96 && dev.L != 128 && dev.L != 256)) {
    set_error("the BF16 filter supports series of length 96, 128 and 256");
    return HKG_EINVAL;
  }
  HKG_HIP_TRY(hipSetDevice(dev.device));
  const uint64_t n = dev.n;
  // rows padded to whole filter tiles (zeros), so that tile copies need no
  // clamp at the end
  const uint64_t ntiles_all = (n + kRowTile - 1) / kRowTile, npad = ntiles_all * kRowTile;
  if (hipMalloc((void **)&dev.Xb, npad * dev.L * 2) != hipSuccess) {
    (void)hipGetLastError();
    dev.Xb = nullptr;
    set_error("no device memory for the BF16 copy of the series");
    return HKG_ENOMEM;
  }
  if (hipMalloc((void **)&dev.xst, npad * sizeof(float4)) != hipSuccess ||
      hipMalloc((void **)&dev.xn2, npad * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&dev.tile_slack, ntiles_all * sizeof(float2)) != hipSuccess) {
    (void)hipGetLastError();
    set_error("no device memory for the BF16 row norms");
    return HKG_ENOMEM;  // freed by the next attempt or the index
  }
  dev.device_bytes += npad * dev.L * 2 + npad * sizeof(float4) + npad * sizeof(float) + ntiles_all * sizeof(float2);
  if (npad > n) {
    HKG_HIP_TRY(hipMemsetAsync(dev.Xb + n * dev.L, 0, (npad - n) * dev.L * 2, stream));
    HKG_HIP_TRY(hipMemsetAsync(dev.xst + n, 0, (npad - n) * sizeof(float4), stream));
    HKG_HIP_TRY(hipMemsetAsync(dev.xn2 + n, 0, (npad - n) * sizeof(float), stream));
  }
  hipLaunchKernelGGL(k_make_bf16, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, n, dev.L, dev.X, dev.Xb,
                     dev.xst, dev.xn2);
  hipLaunchKernelGGL(k_tile_slack, dim3((uint32_t)((ntiles_all + 255) / 256)), dim3(256), 0, stream, n, dev.xst,
                     dev.tile_slack);
  HKG_HIP_TRY(hipGetLastError());
  HKG_HIP_TRY(hipStreamSynchronize(stream));
  if (dev.has_tree) {
    // first and last leaf overlapping every row tile of the filter
    const uint32_t ntiles = (uint32_t)((n + kRowTile - 1) / kRowTile);
    std::vector<uint32_t> l0(ntiles), l1(ntiles);
    const HostTree &t = dev.tree;
    uint32_t leaf = 0;
    for (uint32_t i = 0; i < ntiles; ++i) {
      uint64_t r0 = (uint64_t)i * kRowTile, r1 = std::min<uint64_t>(r0 + kRowTile, n) - 1;
      while (leaf + 1 < dev.nleaf && t.nodes[t.leaves[leaf]].file_pos + t.nodes[t.leaves[leaf]].size <= r0) ++leaf;
      l0[i] = leaf;
      uint32_t e = leaf;
      while (e + 1 < dev.nleaf && t.nodes[t.leaves[e]].file_pos + t.nodes[t.leaves[e]].size <= r1) ++e;
      l1[i] = e;
    }
    if (hipMalloc((void **)&dev.tile_leaf0, ntiles * 4) != hipSuccess ||
        hipMalloc((void **)&dev.tile_leaf1, ntiles * 4) != hipSuccess) {
      (void)hipGetLastError();
      set_error("no device memory for the tile tables");
      return HKG_ENOMEM;
    }
    dev.ntiles = ntiles;
    dev.device_bytes += 8ull * ntiles;
    HKG_HIP_TRY(hipMemcpy(dev.tile_leaf0, l0.data(), ntiles * 4, hipMemcpyHostToDevice));
    HKG_HIP_TRY(hipMemcpy(dev.tile_leaf1, l1.data(), ntiles * 4, hipMemcpyHostToDevice));
  }
  bf16_ready = true;
  return HKG_OK;
}

// filter wave slots of a batch of nq queries (at the smallest slot, 64
// queries) and the 32-bit words of their per-tile need bits
// 32-query slots of a batch (rows of the per-slot leaf needs), and 32-bit
// words of filter wave slots (at least 64 queries each)
static uint32_t bf16_max_slots(uint32_t nq) { return (nq + 31) / 32; }
static uint32_t bf16_slot_words(uint32_t nq) { return ((nq + 63) / 64 + 31) / 32; }

// leaves some query of the batch needs: OR over the 32-query slots
__global__ void k_leaf_need_any(uint32_t nleaf, uint32_t nslot32, const uint8_t *__restrict__ need8,
                                uint8_t *__restrict__ need)
{
  const uint32_t lf = blockIdx.x * blockDim.x + threadIdx.x;
  if (lf >= nleaf) return;
  uint8_t any = 0;
  for (uint32_t s = 0; s < nslot32 && !any; ++s) any = need8[(uint64_t)s * nleaf + lf];
  need[lf] = any ? 1 : 0;
}

// Per row tile, one bit per filter wave slot (sub consecutive 32-query
// slots of the batch order): set when a query of the slot needs a leaf the
// tile overlaps (need8, per 32-query slot and leaf, from k_group_leaf_need)
__global__ void k_tile_slot_bits(uint32_t ntiles, uint32_t nslots, uint32_t words, uint32_t nleaf, uint32_t sub,
                                 uint32_t nslot32, const uint32_t *__restrict__ tile_leaf0,
                                 const uint32_t *__restrict__ tile_leaf1, const uint8_t *__restrict__ need8,
                                 uint32_t *__restrict__ bits)
{
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const uint32_t l0 = tile_leaf0[t], l1 = tile_leaf1[t];
  for (uint32_t w = 0; w < words; ++w) {
    uint32_t v = 0;
    for (uint32_t b = 0; b < 32 && w * 32 + b < nslots; ++b) {
      const uint32_t s0 = (w * 32 + b) * sub, s1 = min(s0 + sub, nslot32);
      bool any = false;
      for (uint32_t s = s0; s < s1 && !any; ++s) {
        const uint8_t *nd = need8 + (uint64_t)s * nleaf;
        for (uint32_t lf = l0; lf <= l1 && !any; ++lf) any = nd[lf] != 0;
      }
      v |= any ? 1u << b : 0u;
    }
    bits[(uint64_t)t * words + w] = v;
  }
}

// The BF16 scratch of a buffer set (survivor lists at the call's cap,
// BF16 queries, block-group tables, survivor region), allocated on first use
hkg_status Index::ensure_bf16_scratch(QueryBuffers &bs)
{
  QueryBuffers *b = &bs;
  const uint32_t L = dev.L, nleaf = dev.nleaf;
  const uint32_t cap_override = call_settings().cap_override;
  const uint32_t want_cap = cap_override ? std::max<uint32_t>(cap_override, 16) : kBf16SurvivorCap;
  // sized for the smallest block group, whatever the shape of a call
  const uint32_t qpb_min = bf16_qpb(L, 2);
  const uint32_t groups_max = (b->B + qpb_min - 1) / qpb_min;
  if (b->bf_surv && b->bf_cap != want_cap) b->free_bf16();
  if (!b->bf_surv) {
    b->bf_cap = want_cap;
    b->bf_sort_bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, b->bf_sort_bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                             (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)b->B, 0, 32);
    if (hipMalloc((void **)&b->bf_surv, (size_t)b->B * b->bf_cap * sizeof(*b->bf_surv)) != hipSuccess ||
        hipMalloc((void **)&b->bf_cnt, (size_t)b->B * 4) != hipSuccess ||
        hipMalloc((void **)&b->Qb, (size_t)b->B * L * 2) != hipSuccess ||
        hipMalloc((void **)&b->qst, (size_t)b->B * sizeof(float4)) != hipSuccess ||
        hipMalloc((void **)&b->qperm, (size_t)b->B * 4 * 4) != hipSuccess ||
        hipMalloc(&b->bf_sort_tmp, std::max<size_t>(b->bf_sort_bytes, 16)) != hipSuccess ||
        hipMalloc((void **)&b->need, (size_t)groups_max * nleaf) != hipSuccess ||
        hipMalloc((void **)&b->tiles, (size_t)groups_max * dev.ntiles * 4) != hipSuccess ||
        hipMalloc((void **)&b->ntiles_g, (size_t)groups_max * 4) != hipSuccess ||
        hipMalloc((void **)&b->thr, (size_t)b->B * 4) != hipSuccess ||
        hipMalloc((void **)&b->hist, (size_t)b->B * kBf16Bins * 4) != hipSuccess ||
        hipMalloc((void **)&b->bf_region, (size_t)bf16_max_blocks() * kRegionCap * sizeof(uint4)) != hipSuccess ||
        hipMalloc((void **)&b->bf_region_cnt, (size_t)bf16_max_blocks() * 4 * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void **)&b->bf_prog, (size_t)bf16_max_blocks() * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void **)&b->slot_need8, (size_t)bf16_max_slots(b->B) * nleaf) != hipSuccess ||
        hipMalloc((void **)&b->slot_need, (size_t)dev.ntiles * bf16_slot_words(b->B) * 4) != hipSuccess) {
      (void)hipGetLastError();
      set_error("no device memory for the BF16 path scratch");
      return HKG_ENOMEM;
    }
  }
  return HKG_OK;
}

hkg_status Index::run_bf16_batch(const float *dQ, const uint32_t *dk, uint32_t nq, uint32_t kmax, hipStream_t st,
                                 QueryBuffers **used, Counters *ctr, Prof *pf, const uint32_t *seed_host,
                                 bool bound_pass, bool use_graph)
{
  if (!dev.has_tree) {
    set_error("the BF16 path needs an index (its first bound comes from the best leaf)");
    return HKG_EINVAL;
  }
  const uint32_t L = dev.L, nleaf = dev.nleaf;
  // the approximate stage's candidates: the per-query-path size, which does
  // not depend on the batch (a server's batches of different sizes then
  // share buffer sets)
  const uint32_t cap = cand_cap(HKG_PATH_INDEX, nq, kmax);
  QueryBuffers *b = nullptr;
  hkg_status rc = get_buffers(nq, kmax, cap, &b);
  if (rc != HKG_OK) return rc;
  BufLease lease{this, b};  // handed to the caller on success
  const uint32_t qt = bf16_query_tiles_ok(L, call_settings().bf16_qt) ? call_settings().bf16_qt : kBf16QueryTiles;
  rc = ensure_bf16_scratch(*b);
  if (rc != HKG_OK) return rc;
  HKG_HIP_TRY(hipMemcpyAsync(b->Q, dQ, (size_t)nq * L * 4, hipMemcpyDeviceToDevice, st));
  HKG_HIP_TRY(hipMemcpyAsync(b->kq, dk, (size_t)nq * 4, hipMemcpyDeviceToDevice, st));
  uint32_t *overflow = b->misc + 2;
  if (seed_host) HKG_HIP_TRY(hipMemcpyAsync(b->seed, seed_host, (size_t)nq * 4, hipMemcpyHostToDevice, st));
  // row tiles the batch needs under the current bounds (one list that
  // every block group walks), and the filter threshold of every query
  float *thr0 = reinterpret_cast<float *>(b->thr);
  constexpr bool kSharedList = true;
  const uint32_t lists = 1;
  // the first pass's tile count is kept on the device (misc[4]) for the
  // counters read after the batch: one list
  static_assert(kSharedList, "the pass-1 tile count has one slot");
  // the whole stage sequence, eager or captured (no host reads inside)
  auto enqueue = [&](Prof *p) -> hkg_status {
    launch_reset(st, nq, b->bound, b->cand_cnt, seed_host ? b->seed : nullptr);
    launch_fill(st, b->misc, 8, 0);
    launch_check_k(st, nq, kmax, b->kq, b->misc + 3);
    launch_fill(st, reinterpret_cast<uint32_t *>(b->counters), 2 * kNumCounters, 0);
    launch_fill(st, b->bf_cnt, nq, 0);
    if (p) p->mark(st, -1);
    launch_prep(st, b->Q, nq, L, b->paa, b->psum, b->qn2);
    launch_query_bf16(st, nq, L, b->Q, b->Qb, b->qst);
    if (p) p->mark(st, HKG_STAGE_PREP);
    launch_eapca(st, nq, nleaf, L, dev.leaf_seg_off, dev.seg_end, dev.seg_ind, b->psum, b->lb);
    if (p) p->mark(st, HKG_STAGE_EAPCA);
    launch_best_leaf(st, nq, nleaf, b->lb, b->best);
    if (p) p->mark(st, HKG_STAGE_SELECT_LEAF);
    launch_scan_best_bf16(st, nq, dev.X, dev.Xb, dev.xn2, dev.max_leaf, L, b->Q, b->kq, dev.sax, b->paa, b->best, dev.leaf_pos,
                          dev.leaf_size, dev.ids, b->cap, b->cand, b->cand_cnt, b->bound, b->counters);
    if (p) p->mark(st, HKG_STAGE_APPROX);
    uint32_t *skeys = b->qperm + b->B, *sidx = b->qperm + 2 * b->B;
    hipLaunchKernelGGL(k_iota_best, dim3((nq + 255) / 256), dim3(256), 0, st, nq, b->best, skeys, sidx);
    HKG_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(b->bf_sort_tmp, b->bf_sort_bytes, skeys, b->qperm + 3 * b->B,
                                                   sidx, b->qperm, (int)nq, 0, 32, st));
    // the leaves each 32-query slot of the batch order needs (the last plan
    // also counts the (query, leaf) pairs it keeps), the leaves any query
    // needs, and the row tiles overlapping those
    const uint32_t nslot32 = bf16_max_slots(nq);
    auto plan_tiles = [&](bool count) {
      launch_fill(st, b->ntiles_g, lists, 0);
      hipLaunchKernelGGL(k_group_leaf_need, dim3((nleaf + 255) / 256, nslot32), dim3(256), 0, st, nq, 32u, nleaf,
                         b->lb, b->bound, b->qperm, b->slot_need8, count ? b->counters + kCtrLeafPairs : nullptr);
      hipLaunchKernelGGL(k_leaf_need_any, dim3((nleaf + 255) / 256), dim3(256), 0, st, nleaf, nslot32, b->slot_need8,
                         b->need);
      hipLaunchKernelGGL(k_group_tiles, dim3((dev.ntiles + 255) / 256, lists), dim3(256), 0, st, dev.ntiles, nleaf,
                         dev.tile_leaf0, dev.tile_leaf1, b->need, b->tiles, b->ntiles_g);
      hipLaunchKernelGGL(k_bf16_thr_init, dim3((nq + 255) / 256), dim3(256), 0, st, nq, b->bound, thr0);
    };
    plan_tiles(!bound_pass);
    if (p) p->mark(st, HKG_STAGE_COLLECT);
    if (bound_pass) {
      // pass 1: histogram of the BF16 upper bounds -> a k-th distance bound
      // within about one bin of the k-th smallest upper bound
      launch_fill(st, b->hist, nq * kBf16Bins, 0);
      launch_bf16_filter(st, dev, nq, b->Qb, b->qst, b->bound, b->qperm, b->tiles, b->ntiles_g, kSharedList,
                         b->bf_cap, b->bf_surv, b->bf_cnt, thr0, b->hist, b->bf_region, b->bf_region_cnt, b->bf_prog,
                         qt, nullptr, 0, nullptr);
      HKG_HIP_TRY(hipMemcpyAsync(b->misc + 4, b->ntiles_g, lists * 4, hipMemcpyDeviceToDevice, st));
      hipLaunchKernelGGL(k_bf16_tau, dim3((nq + 255) / 256), dim3(256), 0, st, nq, b->kq, b->hist, thr0, b->bound);
      plan_tiles(true);
      if (p) p->mark(st, HKG_STAGE_BOUND);
    }
    // the filter wave slots (32 qt consecutive queries of the batch order,
    // qt of the 32-query slots of the last plan) that need each row tile: a
    // wave whose queries need none of the tile's leaves skips its MFMAs
    // (those series lie beyond every one of its queries' k-th distance bound)
    const uint32_t qpw = 32 * qt, nslots = (nq + qpw - 1) / qpw, swords = (nslots + 31) / 32;
    hipLaunchKernelGGL(k_tile_slot_bits, dim3((dev.ntiles + 255) / 256), dim3(256), 0, st, dev.ntiles, nslots, swords,
                       nleaf, qt, nslot32, dev.tile_leaf0, dev.tile_leaf1, b->slot_need8, b->slot_need);
    // pass 2: collect the survivors of the certified lower bound
    launch_bf16_filter(st, dev, nq, b->Qb, b->qst, b->bound, b->qperm, b->tiles, b->ntiles_g, kSharedList, b->bf_cap,
                       b->bf_surv, b->bf_cnt, thr0, nullptr, b->bf_region, b->bf_region_cnt, b->bf_prog, qt,
                       b->slot_need, swords, b->counters);
    if (p) p->mark(st, HKG_STAGE_REFINE);
    launch_bf16_refine(st, nq, dev.X, L, b->Q, b->kq, b->kmax, dev.ids, b->cap, b->cand, b->cand_cnt, b->bf_cap,
                       b->bf_surv, b->bf_cnt, thr0, b->bound, idmap64, b->out_ids, b->out_d2, overflow, b->counters);
    if (p) p->mark(st, HKG_STAGE_TOPK);
    return HKG_OK;
  };
  if (use_graph && !seed_host)
    rc = run_captured(*b, 2, nq, (bound_pass ? 1u : 0u) | qt << 4, st, pf, enqueue);
  else
    rc = enqueue(pf);
  if (rc != HKG_OK) return rc;
  std::vector<uint32_t> nt_pass1;
  if (ctr && bound_pass) {
    nt_pass1.resize(lists);
    HKG_HIP_TRY(hipMemcpyAsync(nt_pass1.data(), b->misc + 4, lists * 4, hipMemcpyDeviceToHost, st));
  }
  HKG_HIP_TRY(hipGetLastError());
  if (ctr) {
    HKG_HIP_TRY(hipMemcpyAsync(ctr->v, b->counters, sizeof(ctr->v), hipMemcpyDeviceToHost, st));
    HKG_HIP_TRY(hipMemcpyAsync(&ctr->overflow, overflow, 4, hipMemcpyDeviceToHost, st));
    HKG_HIP_TRY(hipMemcpyAsync(&ctr->bad_k, b->misc + 3, 4, hipMemcpyDeviceToHost, st));
    std::vector<uint32_t> nt(lists);
    HKG_HIP_TRY(hipMemcpyAsync(nt.data(), b->ntiles_g, lists * 4, hipMemcpyDeviceToHost, st));
    HKG_HIP_TRY(hipStreamSynchronize(st));
    // (query, series) pairs and row tiles the filter passes went through
    // (rows counted once per block group that streamed them)
    const uint64_t groups = (nq + bf16_qpb(L, qt) - 1) / bf16_qpb(L, qt);
    const uint64_t t1 = nt_pass1.empty() ? 0 : nt_pass1[0];
    ctr->v[kCtrBf16] += (unsigned long long)(nt[0] + t1) * kRowTile * nq;
    ctr->bf16_rows += (uint64_t)nt[0] * kRowTile * groups;
    ctr->bf16_rows_bound += t1 * kRowTile * groups;
  }
  *used = b;
  lease.b = nullptr;
  return HKG_OK;
}

// Queries whose BF16 survivor list overflowed run the BF16 pass again,
// seeded with the exact k-th distance of the survivors that were kept (a
// valid upper bound, far tighter than the first one).  Queries still
// overflowing after a few rounds are answered by the FP32 index path.
hkg_status Index::fix_bf16_overflow(QueryBuffers &b, uint32_t nq, uint32_t kmax, hipStream_t st, uint32_t *retries)
{
  const uint32_t L = dev.L;
  std::vector<uint32_t> cnt(nq), kv(nq);
  std::vector<float> d2((size_t)nq * kmax);
  HKG_HIP_TRY(hipMemcpyAsync(cnt.data(), b.bf_cnt, nq * 4, hipMemcpyDeviceToHost, st));
  HKG_HIP_TRY(hipMemcpyAsync(kv.data(), b.kq, nq * 4, hipMemcpyDeviceToHost, st));
  HKG_HIP_TRY(hipMemcpyAsync(d2.data(), b.out_d2, d2.size() * 4, hipMemcpyDeviceToHost, st));
  HKG_HIP_TRY(hipStreamSynchronize(st));
  std::vector<uint32_t> todo;
  for (uint32_t q = 0; q < nq; ++q)
    if (cnt[q] > b.bf_cap) todo.push_back(q);
  if (todo.empty()) return HKG_OK;
  const bool dbg = getenv("HKG_DEBUG") != nullptr;
  std::vector<uint32_t> dbg_bound;
  if (dbg) {
    dbg_bound.resize(nq);
    HKG_HIP_TRY(hipMemcpy(dbg_bound.data(), b.bound, nq * 4, hipMemcpyDeviceToHost));
  }
  uint32_t m = (uint32_t)todo.size();
  float *tq = nullptr;
  uint32_t *tk = nullptr;
  if (hipMalloc((void **)&tq, (size_t)m * L * 4) != hipSuccess || hipMalloc((void **)&tk, (size_t)m * 4) != hipSuccess) {
    (void)hipGetLastError();
    if (tq) (void)hipFree(tq);
    set_error("no device memory for the BF16 overflow re-run");
    return HKG_ENOMEM;
  }
  std::vector<uint32_t> seed(m);
  for (uint32_t i = 0; i < m; ++i) {
    (void)hipMemcpyAsync(tq + (uint64_t)i * L, b.Q + (uint64_t)todo[i] * L, L * 4, hipMemcpyDeviceToDevice, st);
    (void)hipMemcpyAsync(tk + i, b.kq + todo[i], 4, hipMemcpyDeviceToDevice, st);
    float kth = d2[(size_t)todo[i] * kmax + kv[todo[i]] - 1];
    uint32_t bits;
    memcpy(&bits, &kth, 4);
    seed[i] = kth < FLT_MAX ? bits : kInfBits;
  }
  hkg_status rc = HKG_OK;
  const std::vector<uint32_t> todo0 = todo;
  if (dbg) {
    uint32_t with_seed = 0;
    for (uint32_t i = 0; i < m; ++i) with_seed += seed[i] < kInfBits;
    fprintf(stderr, "hkg: bf16 re-run: %u flagged queries (%u with a k-th distance seed)\n", m, with_seed);
  }
  for (int round = 0; round < 4 && m > 0 && rc == HKG_OK; ++round) {
    ++*retries;
    const auto t_round = std::chrono::steady_clock::now();
    QueryBuffers *sb = nullptr;
    BufLease lease{this, nullptr};
    Counters c;
    std::vector<uint32_t> c2(m);
    std::vector<float> o2((size_t)m * kmax);
    if (round < 3) {
      // the first re-run adds the upper-bound histogram pass (unless the
      // caller turned it off); later ones hold exact seeds
      rc = run_bf16_batch(tq, tk, m, kmax, st, &sb, &c, nullptr, seed.data(),
                          round == 0 && call_settings().bound_pass != 2);
      if (rc != HKG_OK) break;
      lease.b = sb;
      if (hipMemcpyAsync(c2.data(), sb->bf_cnt, m * 4, hipMemcpyDeviceToHost, st) != hipSuccess) rc = HKG_EDEVICE;
    } else {
      // last resort: the FP32 index path (its own overflow handling included)
      float *td2 = nullptr;
      uint64_t *tids = nullptr;
      if (hipMalloc((void **)&tids, (size_t)m * kmax * 8) != hipSuccess ||
          hipMalloc((void **)&td2, (size_t)m * kmax * 4) != hipSuccess) {
        (void)hipGetLastError();
        if (tids) (void)hipFree(tids);
        rc = HKG_ENOMEM;
        break;
      }
      hkg_query_opts o;
      hkg_default_query_opts(&o);
      o.path = HKG_PATH_BATCHED;
      const auto t_fb = std::chrono::steady_clock::now();
      rc = knn_device(tq, m, tk, kmax, &o, tids, td2, st, nullptr);
      if (dbg)
        fprintf(stderr, "hkg: bf16 re-run: %u queries on the FP32 path, %.1f ms\n", m,
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_fb).count());
      for (uint32_t i = 0; i < m && rc == HKG_OK; ++i) {
        (void)hipMemcpyAsync(b.out_ids + (uint64_t)todo[i] * kmax, tids + (uint64_t)i * kmax, kmax * 8,
                             hipMemcpyDeviceToDevice, st);
        (void)hipMemcpyAsync(b.out_d2 + (uint64_t)todo[i] * kmax, td2 + (uint64_t)i * kmax, kmax * 4,
                             hipMemcpyDeviceToDevice, st);
      }
      (void)hipStreamSynchronize(st);
      (void)hipFree(tids);
      (void)hipFree(td2);
      break;
    }
    if (hipMemcpyAsync(o2.data(), sb->out_d2, o2.size() * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = HKG_EDEVICE;
      break;
    }
    std::vector<uint32_t> again;
    for (uint32_t i = 0; i < m; ++i) {
      if (c2[i] > sb->bf_cap) {
        float kth = o2[(size_t)i * kmax + kv[todo[i]] - 1];
        uint32_t bits;
        memcpy(&bits, &kth, 4);
        if (kth < FLT_MAX && bits < seed[i]) seed[i] = bits;
        again.push_back(i);
        continue;
      }
      (void)hipMemcpyAsync(b.out_ids + (uint64_t)todo[i] * kmax, sb->out_ids + (uint64_t)i * kmax, kmax * 8,
                           hipMemcpyDeviceToDevice, st);
      (void)hipMemcpyAsync(b.out_d2 + (uint64_t)todo[i] * kmax, sb->out_d2 + (uint64_t)i * kmax, kmax * 4,
                           hipMemcpyDeviceToDevice, st);
    }
    for (uint32_t j = 0; j < again.size(); ++j) {
      uint32_t i = again[j];
      if (i != j) {
        (void)hipMemcpyAsync(tq + (uint64_t)j * L, tq + (uint64_t)i * L, L * 4, hipMemcpyDeviceToDevice, st);
        (void)hipMemcpyAsync(tk + j, tk + i, 4, hipMemcpyDeviceToDevice, st);
      }
      seed[j] = seed[i];
      todo[j] = todo[i];
    }
    if (hipStreamSynchronize(st) != hipSuccess) rc = HKG_EDEVICE;
    if (dbg)
      fprintf(stderr, "hkg: bf16 re-run round %d: %u queries, %u again, %.1f ms\n", round, m, (uint32_t)again.size(),
              std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_round).count());
    m = (uint32_t)again.size();
  }
  (void)hipFree(tq);
  (void)hipFree(tk);
  if (dbg && rc == HKG_OK) {
    // first flagged queries: bound, k-th found by the flagged pass, final k-th
    std::vector<float> fin((size_t)nq * kmax);
    if (hipMemcpyAsync(fin.data(), b.out_d2, fin.size() * 4, hipMemcpyDeviceToHost, st) == hipSuccess &&
        hipStreamSynchronize(st) == hipSuccess) {
      for (size_t i = 0; i < todo0.size() && i < 12; ++i) {
        const uint32_t q = todo0[i];
        float bnd;
        memcpy(&bnd, &dbg_bound[q], 4);
        fprintf(stderr, "hkg:   q %u k %u bound %.6g found %.6g final %.6g\n", q, kv[q], bnd,
                d2[(size_t)q * kmax + kv[q] - 1], fin[(size_t)q * kmax + kv[q] - 1]);
      }
    }
  }
  return rc;
}

void launch_query_bf16(hipStream_t st, uint32_t nq, uint32_t L, const float *Q, uint16_t *Qb, float4 *qst)
{
  hipLaunchKernelGGL(k_query_bf16, dim3((nq + 63) / 64), dim3(64), 0, st, nq, L, Q, Qb, qst);
}

}  // namespace hkg
