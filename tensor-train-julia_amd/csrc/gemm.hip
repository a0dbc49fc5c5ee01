// Grouped (variable-size, multi-segment) GEMM on CDNA4 MFMA.
//
// One workgroup computes one 64x64 tile of one problem.  The tile index of a
// workgroup is mapped to its problem by a binary search over the problems'
// first-tile prefix, so one launch covers every (item, term, block) product of
// a sweep step, whatever their shapes.  Operands are staged through LDS with
// the loader choosing the coalesced direction from the strides; A-operand
// values are converted to the math precision (and scaled per segment) on load.
//
// Math precisions:
//   F64  : v_mfma_f64_16x16x4_f64, FP64 accumulate
//   F32  : v_mfma_f32_16x16x4_f32, FP32 accumulate
//   BF16 : v_mfma_f32_16x16x32_bf16, FP32 accumulate
#include <cmath>
#include <cstring>

#include "ttr_internal.h"

namespace ttr {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

struct bf16s {
  uint16_t x;
};

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40);  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

template <class T> __device__ __forceinline__ double load_d(const void* p, int64_t i);
template <> __device__ __forceinline__ double load_d<double>(const void* p, int64_t i) { return static_cast<const double*>(p)[i]; }
template <> __device__ __forceinline__ double load_d<float>(const void* p, int64_t i) { return static_cast<const float*>(p)[i]; }
template <> __device__ __forceinline__ double load_d<bf16s>(const void* p, int64_t i) { return bf16_to_f32(static_cast<const uint16_t*>(p)[i]); }

// operands live in global memory: global (not flat) loads
template <class T> using gptr = const T __attribute__((address_space(1)))*;
template <class T> __device__ __forceinline__ T ldg(const void* p, int64_t i) { return ((gptr<T>)p)[i]; }
template <> __device__ __forceinline__ bf16s ldg<bf16s>(const void* p, int64_t i) {
  bf16s r;
  r.x = ((gptr<uint16_t>)p)[i];
  return r;
}

__device__ __forceinline__ double to_double(double v) { return v; }
__device__ __forceinline__ double to_double(float v) { return v; }
__device__ __forceinline__ double to_double(bf16s v) { return bf16_to_f32(v.x); }

template <class T> __device__ __forceinline__ void store_c(void* p, int64_t i, double v);
template <> __device__ __forceinline__ void store_c<double>(void* p, int64_t i, double v) { static_cast<double*>(p)[i] = v; }
template <> __device__ __forceinline__ void store_c<float>(void* p, int64_t i, double v) { static_cast<float*>(p)[i] = float(v); }
template <> __device__ __forceinline__ void store_c<bf16s>(void* p, int64_t i, double v) { static_cast<uint16_t*>(p)[i] = f32_to_bf16(float(v)); }

// The descriptors of a small batch travel in the kernel arguments (the kernel
// argument segment holds up to 4 KiB), so such a launch needs no upload: one
// copy dispatch and one dependent launch boundary less.  Larger batches, and
// launches recorded into a graph (whose replays rewrite uploaded descriptors),
// pass device pointers.
constexpr int GEMM_INL_PROBS = 8, GEMM_INL_SEGS = 16;
template <bool INL> struct GemmInl {};
template <> struct GemmInl<true> {
  GemmProb p[GEMM_INL_PROBS];
  GemmSeg s[GEMM_INL_SEGS];
  int32_t ids[GEMM_INL_PROBS];  // the problems whose K is split (for the reduction)
};
static_assert(sizeof(GemmInl<true>) <= 3072, "GEMM descriptors fit in the kernel arguments");

__device__ __forceinline__ int find_prob(const GemmProb* probs, int nprobs, int bid) {
  int lo = 0, hi = nprobs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (probs[mid].tile0 <= bid) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ double amode_value(int mode, int64_t gi, int64_t gl, double v) {
  if (mode == AM_VT_UNIT) return gl < gi ? 0.0 : (gl == gi ? 1.0 : v);
  if (mode == AM_V_UNIT) return gi < gl ? 0.0 : (gi == gl ? 1.0 : v);
  return v;
}

constexpr int BM = 64, BN = 64;  // tile of the BF16 kernel

// ------------------------------------------------------------- F64 / F32 math
// Tile TM x TN (64 or 128 each), k-step BK (8 or 16), 256 threads as 2 x 2 waves;
// wave (wm, wn) owns a (TM/2) x (TN/2) block made of (TM/32) x (TN/32) MFMA 16x16
// outputs.  The accumulation order of every output element is the K order whatever
// the tile, and the split-K chunks are counted in k-tiles of KS = 16 for every BK,
// so all tile configurations give bitwise identical results.
//
// Pipeline: two LDS buffers.  While the MFMAs of k-step i read one buffer, the
// global loads of k-step i + 1 are in flight into registers and are then stored
// to the other buffer, so one barrier per k-step orders everything.  The k-steps
// of all segments of a problem form one sequence, so the pipeline does not drain
// at segment boundaries.
constexpr int KS = 16;
constexpr int NXCD = 8;  // accelerator complex dies; workgroups go to them round-robin

// Workgroup bid runs on die bid % 8.  Tile index given to it: each die gets one
// contiguous range of the tile sequence, so the tiles that share operand panels
// (neighbouring tiles of one problem, and the problems of one item) meet in the
// same die's L2 instead of being spread over all eight.
__device__ __forceinline__ int xcd_tile(int bid, int n) {
  const int q = n / NXCD, r = n - q * NXCD;
  const int x = bid % NXCD, j = bid / NXCD;
  return x * q + min(x, r) + j;
}

template <class TA, class TB, class TC, class MT, int TM, int TN, int BK, bool INL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TM * TN >= 128 * 128 ? 2 : 1)))
void gemm_scalar_kernel(const GemmProb* __restrict__ probs, int nprobs, const GemmSeg* __restrict__ segs,
                        GemmInl<INL> inl) {
  static_assert(KS % BK == 0, "a split-K k-tile holds whole k-steps");
  if constexpr (INL) {
    probs = inl.p;
    segs = inl.s;
  }
  constexpr int MI = TM / 32, NI = TN / 32;
  constexpr int NA = TM * BK / 256, NB = TN * BK / 256;  // tile elements per thread
  __shared__ MT As[2][BK][TM + 1];
  __shared__ MT Bs[2][BK][TN + 1];
  using acc_t = typename std::conditional<std::is_same<MT, double>::value, d4_t, f4_t>::type;

  const int bid = xcd_tile(blockIdx.x, gridDim.x);
  const GemmProb P = probs[find_prob(probs, nprobs, bid)];
  const int t0 = bid - P.tile0;
  const int chunk = t0 / P.tiles_mn, t = t0 - chunk * P.tiles_mn;
  const int row0 = (t / P.tiles_n) * TM, col0 = (t % P.tiles_n) * TN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  // k-tiles (of KS) [kt_lo, kt_hi) of the problem's segment sequence belong to this workgroup
  const int kt_lo = P.ksplit > 1 ? chunk * P.kc : 0;
  const int kt_hi = P.ksplit > 1 ? kt_lo + P.kc : (1 << 30);

  acc_t acc[MI][NI];
#pragma unroll
  for (int a = 0; a < MI; ++a)
#pragma unroll
    for (int b = 0; b < NI; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

  int emax = -(1 << 30);
  for (int s = 0; s < P.nseg; ++s)
    if (segs[P.seg0 + s].exp) emax = max(emax, segs[P.seg0 + s].exp[0]);

  // Element q of a thread's share of the A tile is (i, l) = (ai0, al0 + q * adl)
  // when A is column-major in memory (i fastest across threads), else
  // (ai0 + q * adi, al0) (l fastest); likewise (l, j) for B.  The thread's first
  // element and the step between its elements are fixed per segment.
  TA ra[NA];
  TB rb[NB];
  int buf = 0;
  bool started = false;
  int kt_seg = 0;  // first k-tile of the segment in the problem's sequence
  for (int s = 0; s < P.nseg; ++s) {
    const GemmSeg S = segs[P.seg0 + s];
    const int nkt = (S.K + KS - 1) / KS;
    const int kb = max(kt_lo - kt_seg, 0) * KS, ke = min(kt_hi - kt_seg, nkt) * KS;
    kt_seg += nkt;
    if (kb >= ke) continue;
    const int kend = min(ke, S.K);
    double scale = S.scale;
    if (S.exp) scale = ldexp(scale, S.exp[0] - emax);
    int ai0, al0, adi, adl, bl0, bj0, bdl, bdj;
    if (S.a_rs == 1) { ai0 = tid % TM; al0 = tid / TM; adi = 0; adl = 256 / TM; }
    else { al0 = tid % BK; ai0 = tid / BK; adl = 0; adi = 256 / BK; }
    if (S.b_rs == 1) { bl0 = tid % BK; bj0 = tid / BK; bdl = 0; bdj = 256 / BK; }
    else { bj0 = tid % TN; bl0 = tid / TN; bdj = 0; bdl = 256 / TN; }
    const int64_t aoff = int64_t(row0 + ai0) * S.a_rs + int64_t(al0) * S.a_cs;
    const int64_t adq = int64_t(adi) * S.a_rs + int64_t(adl) * S.a_cs;
    const int64_t boff = int64_t(bl0) * S.b_rs + int64_t(col0 + bj0) * S.b_cs;
    const int64_t bdq = int64_t(bdl) * S.b_rs + int64_t(bdj) * S.b_cs;
    const bool a_in = row0 + TM <= P.M && S.a_mode == AM_PLAIN, b_in = col0 + TN <= P.N;

    // Raw loads of k-step k0 into registers (no arithmetic on the loaded values,
    // so nothing waits for them before the MFMAs); the unit diagonal of the
    // implicit-V modes and the segment scale are applied when the registers are
    // stored to LDS.  Tiles entirely inside the operands load without predicates.
    auto load_tiles = [&](int k0) {
      const bool k_full = k0 + BK <= S.K;
      const int64_t ao = aoff + int64_t(k0) * S.a_cs, bo = boff + int64_t(k0) * S.b_rs;
      if (a_in && k_full) {
#pragma unroll
        for (int q = 0; q < NA; ++q) ra[q] = ldg<TA>(S.A, ao + q * adq);
      } else {
#pragma unroll
        for (int q = 0; q < NA; ++q) {
          const int gi = row0 + ai0 + q * adi, gl = k0 + al0 + q * adl;
          const bool in = gi < P.M && gl < S.K &&
                          (S.a_mode == AM_PLAIN || (S.a_mode == AM_VT_UNIT ? gl > gi : gi > gl));
          ra[q] = TA{};
          if (in) ra[q] = ldg<TA>(S.A, ao + q * adq);
        }
      }
      if (b_in && k_full) {
#pragma unroll
        for (int q = 0; q < NB; ++q) rb[q] = ldg<TB>(S.B, bo + q * bdq);
      } else {
#pragma unroll
        for (int q = 0; q < NB; ++q) {
          const int gl = k0 + bl0 + q * bdl, gj = col0 + bj0 + q * bdj;
          rb[q] = TB{};
          if (gl < S.K && gj < P.N) rb[q] = ldg<TB>(S.B, bo + q * bdq);
        }
      }
    };
    auto store_tiles = [&](int b, int k0) {
      if (S.a_mode == AM_PLAIN) {
#pragma unroll
        for (int q = 0; q < NA; ++q) As[b][al0 + q * adl][ai0 + q * adi] = MT(to_double(ra[q]) * scale);
      } else {
#pragma unroll
        for (int q = 0; q < NA; ++q) {
          const int i = ai0 + q * adi, l = al0 + q * adl;
          double v = to_double(ra[q]);
          if (row0 + i == k0 + l && row0 + i < P.M && k0 + l < S.K) v = 1.0;
          As[b][l][i] = MT(v * scale);
        }
      }
#pragma unroll
      for (int q = 0; q < NB; ++q) Bs[b][bl0 + q * bdl][bj0 + q * bdj] = MT(to_double(rb[q]));
    };

    // the first k-step of a segment goes to the buffer the previous segment's
    // last MFMAs did not read
    if (started) buf ^= 1;
    started = true;
    load_tiles(kb);
    store_tiles(buf, kb);
    __syncthreads();
    for (int k0 = kb; k0 < kend; k0 += BK) {
      const bool more = k0 + BK < kend;
      if (more) load_tiles(k0 + BK);  // in flight during this k-step's MFMAs
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        MT av[MI], bv[NI];
#pragma unroll
        for (int a = 0; a < MI; ++a) av[a] = As[buf][kk + (lane >> 4)][wm * (TM / 2) + 16 * a + (lane & 15)];
#pragma unroll
        for (int b = 0; b < NI; ++b) bv[b] = Bs[buf][kk + (lane >> 4)][wn * (TN / 2) + 16 * b + (lane & 15)];
#pragma unroll
        for (int a = 0; a < MI; ++a)
#pragma unroll
          for (int b = 0; b < NI; ++b) {
            if constexpr (std::is_same<MT, double>::value)
              acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
            else
              acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[b], acc[a][b], 0, 0, 0);
          }
      }
      if (more) {
        store_tiles(buf ^ 1, k0 + BK);
        __syncthreads();
        buf ^= 1;
      }
    }
  }
  // epilogue: register r of a lane holds column lane%16 of each 16x16 block and
  // row 4*(lane/16)+r for the f32 MFMA, row 4*r+lane/16 for the f64 MFMA
#pragma unroll
  for (int bi = 0; bi < MI; ++bi)
#pragma unroll
    for (int bj = 0; bj < NI; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = std::is_same<MT, double>::value ? 4 * r + (lane >> 4) : 4 * (lane >> 4) + r;
        const int64_t gi = row0 + wm * (TM / 2) + bi * 16 + rr;
        const int64_t gj = col0 + wn * (TN / 2) + bj * 16 + (lane & 15);
        if (gi < P.M && gj < P.N) {
          if (P.ksplit > 1) {
            P.part[(int64_t(chunk) * P.N + gj) * P.M + gi] = double(acc[bi][bj][r]);
            continue;
          }
          const int64_t ci = gi * P.c_rs + gj * P.c_cs;
          double v = P.alpha * double(acc[bi][bj][r]);
          if (P.beta != 0.0) v += P.beta * load_d<TC>(P.C, ci);
          store_c<TC>(P.C, ci, v);
        }
      }
}

// C = alpha * (sum of the chunks in chunk order) + beta * C for the split problems
// (blockIdx.y indexes the list of split problems)
template <class TC, bool INL>
__global__ __launch_bounds__(256) void gemm_splitk_reduce(const GemmProb* __restrict__ probs,
                                                          const int32_t* __restrict__ split_ids, GemmInl<INL> inl) {
  if constexpr (INL) {
    probs = inl.p;
    split_ids = inl.ids;
  }
  const GemmProb P = probs[split_ids[blockIdx.y]];
  const int64_t mn = int64_t(P.M) * P.N;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < mn; e += int64_t(gridDim.x) * blockDim.x) {
    double v = 0.0;
    for (int c = 0; c < P.ksplit; ++c) v += P.part[c * mn + e];
    const int64_t gi = e % P.M, gj = e / P.M;
    const int64_t ci = gi * P.c_rs + gj * P.c_cs;
    v *= P.alpha;
    if (P.beta != 0.0) v += P.beta * load_d<TC>(P.C, ci);
    store_c<TC>(P.C, ci, v);
  }
}

// ------------------------------------------------------------- BF16 math
// Tile 64 x 64, k-step 32, 256 threads as 2 x 2 waves of 32 x 32 (2 x 2 MFMA
// 16x16x32 outputs).  LDS holds A as [m][k] and B as [n][k] (k contiguous) so each
// lane reads its 8 consecutive k values of an MFMA operand as one 16-byte vector;
// values are rounded to BF16 when they are stored to LDS (the FP64 core operand
// of the first sketch contraction is read as FP64).  Pipeline as the scalar
// kernel: two LDS buffers, the raw loads of k-step i + 1 in flight during the
// MFMAs of k-step i, one barrier per k-step, unpredicated loads for tiles inside
// the operands, die-contiguous tile order.  K is never split for BF16.
template <class TA, class TB, class TC, bool INL>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(const GemmProb* __restrict__ probs, int nprobs,
                                                        const GemmSeg* __restrict__ segs, GemmInl<INL> inl) {
  constexpr int BK = 32;
  if constexpr (INL) {
    probs = inl.p;
    segs = inl.s;
  }
  constexpr int KP = BK + 8;                              // padded row (elements)
  constexpr int NA = BM * BK / 256, NB = BN * BK / 256;  // tile elements per thread
  __shared__ __attribute__((aligned(16))) uint16_t As[2][BM][KP];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][BN][KP];

  const int bid = xcd_tile(blockIdx.x, gridDim.x);
  const GemmProb P = probs[find_prob(probs, nprobs, bid)];
  const int t = bid - P.tile0;
  const int row0 = (t / P.tiles_n) * BM, col0 = (t % P.tiles_n) * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  f4_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f4_t{0, 0, 0, 0};

  int emax = -(1 << 30);
  for (int s = 0; s < P.nseg; ++s)
    if (segs[P.seg0 + s].exp) emax = max(emax, segs[P.seg0 + s].exp[0]);

  TA ra[NA];
  TB rb[NB];
  int buf = 0;
  bool started = false;
  for (int s = 0; s < P.nseg; ++s) {
    const GemmSeg S = segs[P.seg0 + s];
    if (S.K <= 0) continue;
    double scale = S.scale;
    if (S.exp) scale = ldexp(scale, S.exp[0] - emax);
    // element q of the thread's share: A (i, l) = (ai0 + q adi, al0 + q adl),
    // B (l, j) = (bl0 + q bdl, bj0 + q bdj), the fast index along memory
    int ai0, al0, adi, adl, bl0, bj0, bdl, bdj;
    if (S.a_rs == 1) { ai0 = tid % BM; al0 = tid / BM; adi = 0; adl = 256 / BM; }
    else { al0 = tid % BK; ai0 = tid / BK; adl = 0; adi = 256 / BK; }
    if (S.b_rs == 1) { bl0 = tid % BK; bj0 = tid / BK; bdl = 0; bdj = 256 / BK; }
    else { bj0 = tid % BN; bl0 = tid / BN; bdj = 0; bdl = 256 / BN; }
    const int64_t aoff = int64_t(row0 + ai0) * S.a_rs + int64_t(al0) * S.a_cs;
    const int64_t adq = int64_t(adi) * S.a_rs + int64_t(adl) * S.a_cs;
    const int64_t boff = int64_t(bl0) * S.b_rs + int64_t(col0 + bj0) * S.b_cs;
    const int64_t bdq = int64_t(bdl) * S.b_rs + int64_t(bdj) * S.b_cs;
    const bool a_in = row0 + BM <= P.M && S.a_mode == AM_PLAIN, b_in = col0 + BN <= P.N;
    auto load_tiles = [&](int k0) {
      const bool k_full = k0 + BK <= S.K;
      const int64_t ao = aoff + int64_t(k0) * S.a_cs, bo = boff + int64_t(k0) * S.b_rs;
      if (a_in && k_full) {
#pragma unroll
        for (int q = 0; q < NA; ++q) ra[q] = ldg<TA>(S.A, ao + q * adq);
      } else {
#pragma unroll
        for (int q = 0; q < NA; ++q) {
          const int gi = row0 + ai0 + q * adi, gl = k0 + al0 + q * adl;
          const bool in = gi < P.M && gl < S.K &&
                          (S.a_mode == AM_PLAIN || (S.a_mode == AM_VT_UNIT ? gl > gi : gi > gl));
          ra[q] = TA{};
          if (in) ra[q] = ldg<TA>(S.A, ao + q * adq);
        }
      }
      if (b_in && k_full) {
#pragma unroll
        for (int q = 0; q < NB; ++q) rb[q] = ldg<TB>(S.B, bo + q * bdq);
      } else {
#pragma unroll
        for (int q = 0; q < NB; ++q) {
          const int gl = k0 + bl0 + q * bdl, gj = col0 + bj0 + q * bdj;
          rb[q] = TB{};
          if (gl < S.K && gj < P.N) rb[q] = ldg<TB>(S.B, bo + q * bdq);
        }
      }
    };
    auto store_tiles = [&](int b, int k0) {
#pragma unroll
      for (int q = 0; q < NA; ++q) {
        const int i = ai0 + q * adi, l = al0 + q * adl;
        double v = to_double(ra[q]);
        if (S.a_mode != AM_PLAIN && row0 + i == k0 + l && row0 + i < P.M && k0 + l < S.K) v = 1.0;
        As[b][i][l] = f32_to_bf16(float(v * scale));
      }
#pragma unroll
      for (int q = 0; q < NB; ++q) Bs[b][bj0 + q * bdj][bl0 + q * bdl] = f32_to_bf16(float(to_double(rb[q])));
    };
    if (started) buf ^= 1;
    started = true;
    load_tiles(0);
    store_tiles(buf, 0);
    __syncthreads();
    for (int k0 = 0; k0 < S.K; k0 += BK) {
      const bool more = k0 + BK < S.K;
      if (more) load_tiles(k0 + BK);  // in flight during this k-step's MFMAs
      const int kq = 8 * (lane >> 4);
      const bf8_t a0 = *reinterpret_cast<const bf8_t*>(&As[buf][wm * 32 + (lane & 15)][kq]);
      const bf8_t a1 = *reinterpret_cast<const bf8_t*>(&As[buf][wm * 32 + 16 + (lane & 15)][kq]);
      const bf8_t b0 = *reinterpret_cast<const bf8_t*>(&Bs[buf][wn * 32 + (lane & 15)][kq]);
      const bf8_t b1 = *reinterpret_cast<const bf8_t*>(&Bs[buf][wn * 32 + 16 + (lane & 15)][kq]);
      acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
      if (more) {
        store_tiles(buf ^ 1, k0 + BK);
        __syncthreads();
        buf ^= 1;
      }
    }
  }
  // register r of an f32 16x16 MFMA result holds row 4 (lane / 16) + r, column lane % 16
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = row0 + wm * 32 + bi * 16 + 4 * (lane >> 4) + r;
        const int64_t gj = col0 + wn * 32 + bj * 16 + (lane & 15);
        if (gi < P.M && gj < P.N) {
          const int64_t ci = gi * P.c_rs + gj * P.c_cs;
          double v = P.alpha * double(acc[bi][bj][r]);
          if (P.beta != 0.0) v += P.beta * load_d<TC>(P.C, ci);
          store_c<TC>(P.C, ci, v);
        }
      }
}

// ------------------------------------------------------------- host side
void GemmBatch::prob(void* C, int64_t c_rs, int64_t c_cs, int M, int N, double alpha, double beta) {
  GemmProb p{};
  p.C = C;
  p.c_rs = c_rs;
  p.c_cs = c_cs;
  p.M = M;
  p.N = N;
  p.alpha = alpha;
  p.beta = beta;
  p.seg0 = int32_t(segs_.size());
  p.nseg = 0;
  probs_.push_back(p);
}

void GemmBatch::seg(const void* A, int64_t a_rs, int64_t a_cs, const void* B, int64_t b_rs, int64_t b_cs,
                    int K, double scale, int a_mode, const int32_t* exp) {
  GemmSeg s{};
  s.A = A;
  s.B = B;
  s.a_rs = a_rs;
  s.a_cs = a_cs;
  s.b_rs = b_rs;
  s.b_cs = b_cs;
  s.K = K;
  s.a_mode = a_mode;
  s.scale = scale;
  s.exp = exp;
  segs_.push_back(s);
  probs_.back().nseg += 1;
}

namespace {
// tile configurations of the scalar kernel: {TM, TN, BK}
constexpr int NCFG = 8;
constexpr int CFG_TM[NCFG] = {64, 128, 64, 128, 64, 128, 64, 128};
constexpr int CFG_TN[NCFG] = {64, 64, 128, 128, 64, 64, 128, 128};

template <class TA, class TB, class TC, class MT, bool INL>
void launch_scalar(Ctx& ctx, int cfg, int ntiles, const GemmProb* dp, int np, const GemmSeg* ds,
                   const GemmInl<INL>& inl) {
  dim3 grid(ntiles), block(256);
  switch (cfg) {
    case 0: TTR_LAUNCH(ctx, (gemm_scalar_kernel<TA, TB, TC, MT, 64, 64, 16, INL>), grid, block, 0, dp, np, ds, inl); break;
    case 1: TTR_LAUNCH(ctx, (gemm_scalar_kernel<TA, TB, TC, MT, 128, 64, 16, INL>), grid, block, 0, dp, np, ds, inl); break;
    case 2: TTR_LAUNCH(ctx, (gemm_scalar_kernel<TA, TB, TC, MT, 64, 128, 16, INL>), grid, block, 0, dp, np, ds, inl); break;
    case 3: TTR_LAUNCH(ctx, (gemm_scalar_kernel<TA, TB, TC, MT, 128, 128, 16, INL>), grid, block, 0, dp, np, ds, inl); break;
    case 4: TTR_LAUNCH(ctx, (gemm_scalar_kernel<TA, TB, TC, MT, 64, 64, 8, INL>), grid, block, 0, dp, np, ds, inl); break;
    case 5: TTR_LAUNCH(ctx, (gemm_scalar_kernel<TA, TB, TC, MT, 128, 64, 8, INL>), grid, block, 0, dp, np, ds, inl); break;
    case 6: TTR_LAUNCH(ctx, (gemm_scalar_kernel<TA, TB, TC, MT, 64, 128, 8, INL>), grid, block, 0, dp, np, ds, inl); break;
    default: TTR_LAUNCH(ctx, (gemm_scalar_kernel<TA, TB, TC, MT, 128, 128, 8, INL>), grid, block, 0, dp, np, ds, inl); break;
  }
}

// the precision combinations the library uses: (A, B, C, math)
template <class TA, class TB, class TC, bool INL>
void launch_combo(Ctx& ctx, Prec math, int cfg, int ntiles, const GemmProb* dp, int np, const GemmSeg* ds,
                  const GemmInl<INL>& inl) {
  if (math == Prec::F64) launch_scalar<TA, TB, TC, double>(ctx, cfg, ntiles, dp, np, ds, inl);
  else if (math == Prec::F32) launch_scalar<TA, TB, TC, float>(ctx, cfg, ntiles, dp, np, ds, inl);
  else TTR_LAUNCH(ctx, (gemm_bf16_kernel<TA, TB, TC, INL>), dim3(ntiles), dim3(256), 0, dp, np, ds, inl);
}

template <bool INL>
void launch(Ctx& ctx, Prec ta, Prec tb, Prec tc, Prec math, int cfg, int ntiles, const GemmProb* dp, int np,
            const GemmSeg* ds, const GemmInl<INL>& inl) {
  using P = Prec;
  if (ta == P::F64 && tb == P::F64 && tc == P::F64) launch_combo<double, double, double>(ctx, math, cfg, ntiles, dp, np, ds, inl);
  else if (ta == P::F64 && tb == P::F32 && tc == P::F32) launch_combo<double, float, float>(ctx, math, cfg, ntiles, dp, np, ds, inl);
  else if (ta == P::F64 && tb == P::BF16 && tc == P::BF16) launch_combo<double, bf16s, bf16s>(ctx, math, cfg, ntiles, dp, np, ds, inl);
  else if (ta == P::F32 && tb == P::F32 && tc == P::F32) launch_combo<float, float, float>(ctx, math, cfg, ntiles, dp, np, ds, inl);
  else if (ta == P::BF16 && tb == P::BF16 && tc == P::BF16) launch_combo<bf16s, bf16s, bf16s>(ctx, math, cfg, ntiles, dp, np, ds, inl);
  else if (ta == P::F64 && tb == P::F32 && tc == P::F64) launch_combo<double, float, double>(ctx, math, cfg, ntiles, dp, np, ds, inl);
  else if (ta == P::F64 && tb == P::BF16 && tc == P::F64) launch_combo<double, bf16s, double>(ctx, math, cfg, ntiles, dp, np, ds, inl);
  else throw Error(ST_INVALID_ARG, "unsupported GEMM precision combination");
  TTR_HIP(hipGetLastError());
}

int ilog2(double v) { return v < 1.0 ? 0 : int(std::floor(std::log2(v))); }
}  // namespace

void gemm_run(Ctx& ctx, GemmBatch& b, Prec ta, Prec tb, Prec tc, Prec math, const char* site) {
  std::vector<GemmProb> probs;
  std::vector<GemmSeg> segs = std::move(b.segs_);
  probs.reserve(b.probs_.size());
  double sm = 0.0, sn = 0.0, sk = 0.0, flops = 0.0, bytes = 0.0;
  bool beta0 = true;
  for (const auto& p : b.probs_) {
    if (p.M <= 0 || p.N <= 0) continue;
    probs.push_back(p);
    double k = 0.0;
    for (int q = 0; q < p.nseg; ++q) {
      const GemmSeg& sg = segs[p.seg0 + q];
      k += sg.K;
      bytes += double(p.M) * sg.K * prec_bytes(ta) + double(sg.K) * p.N * prec_bytes(tb);
    }
    sm += p.M;
    sn += p.N;
    sk += k;
    flops += 2.0 * p.M * p.N * k;
    bytes += double(p.M) * p.N * prec_bytes(tc) * (p.beta != 0.0 ? 2.0 : 1.0);
    beta0 = beta0 && p.beta == 0.0;
  }
  b.clear();
  if (probs.empty()) return;
  const int np = int(probs.size());
  const double am = sm / np, an = sn / np, ak = sk / np;
  // tile choice: measured once per (site, precisions, size class) when the batch
  // can be run more than once (beta == 0, C never aliases A or B at any call
  // site), otherwise 64 x 64 (measured fastest for the accumulating updates)
  int cfg = 0;
  bool tune = false;
  std::string key;
  if (math != Prec::BF16) {
    key = std::string(site ? site : "-") + "|" + std::to_string(int(ta)) + std::to_string(int(tb)) +
          std::to_string(int(tc)) + std::to_string(int(math)) + "|" + std::to_string(ilog2(am)) + "," +
          std::to_string(ilog2(an)) + "," + std::to_string(ilog2(ak)) + "," + std::to_string(ilog2(np));
    auto it = ctx.gemm_cfg.find(key);
    if (ctx.gemm_force_cfg >= 0) cfg = ctx.gemm_force_cfg;
    else if (it != ctx.gemm_cfg.end()) cfg = it->second;
    else if (ctx.gemm_autotune && beta0 && !ctx.ws.fixed()) tune = true;
  } else {
    cfg = 0;
  }
  // Split K when the batch has too few 64 x 64 output tiles to fill the chip.
  // The split depends on the shapes only (not on the tile configuration), so
  // every configuration the tuner tries still writes the same C.
  std::vector<int32_t> split_ids;
  for (auto& p : probs) {
    p.ksplit = 1;
    p.kc = 0;
    p.part = nullptr;
  }
  if (math != Prec::BF16) {
    int64_t t64 = 0;
    for (const auto& p : probs) t64 += int64_t((p.M + 63) / 64) * ((p.N + 63) / 64);
    if (t64 < 512) {
      int want = int(std::min<int64_t>(16, (1024 + t64 - 1) / t64));
      for (;; want /= 2) {
        int64_t elems = 0;
        for (auto& p : probs) {
          int kt = 0;
          for (int q = 0; q < p.nseg; ++q) kt += (segs[p.seg0 + q].K + 15) / 16;
          int sp = std::min(want, kt / 8);  // chunks of at least 8 k-tiles
          if (sp <= 1) { p.ksplit = 1; continue; }
          p.kc = (kt + sp - 1) / sp;
          p.ksplit = (kt + p.kc - 1) / p.kc;
          elems += int64_t(p.ksplit) * p.M * p.N;
        }
        if (elems <= (int64_t(1) << 25) || want <= 2) {
          if (elems > (int64_t(1) << 25))
            for (auto& p : probs) p.ksplit = 1;
          else if (elems > 0) {
            double* base = ctx.splitk_buffer(size_t(elems));
            if (!base) {  // no room in a fixed workspace: K is not split
              for (auto& p : probs) p.ksplit = 1;
              break;
            }
            int64_t off = 0;
            for (size_t i = 0; i < probs.size(); ++i) {
              auto& p = probs[i];
              if (p.ksplit <= 1) continue;
              p.part = base + off;
              off += int64_t(p.ksplit) * p.M * p.N;
              split_ids.push_back(int32_t(i));
              bytes += 16.0 * p.ksplit * p.M * p.N;
            }
          }
          break;
        }
      }
    }
  }
  // the tile configurations to run: every admissible one when tuning
  std::vector<int> cands;
  if (tune) {
    for (int c = 0; c < NCFG; ++c)
      if (!((CFG_TM[c] == 128 && am < 96) || (CFG_TN[c] == 128 && an < 96))) cands.push_back(c);
  } else {
    cands.push_back(cfg);
  }
  // one upload for all descriptors: segments, split ids, then the problems as
  // laid out for each candidate configuration
  auto align16 = [](size_t x) { return (x + 15) & ~size_t(15); };
  const size_t seg_bytes = align16(segs.size() * sizeof(GemmSeg));
  const size_t ids_bytes = align16(split_ids.size() * sizeof(int32_t));
  const size_t prob_bytes = align16(probs.size() * sizeof(GemmProb));
  std::vector<char> pack(seg_bytes + ids_bytes + prob_bytes * cands.size());
  std::memcpy(pack.data(), segs.data(), segs.size() * sizeof(GemmSeg));
  if (!split_ids.empty()) std::memcpy(pack.data() + seg_bytes, split_ids.data(), split_ids.size() * sizeof(int32_t));
  std::vector<int64_t> ntiles(cands.size());
  for (size_t ci = 0; ci < cands.size(); ++ci) {
    const int c = cands[ci];
    const int TM = math == Prec::BF16 ? BM : CFG_TM[c], TN = math == Prec::BF16 ? BN : CFG_TN[c];
    int64_t nt = 0;
    for (auto& p : probs) {
      const int tm = (p.M + TM - 1) / TM, tn = (p.N + TN - 1) / TN;
      p.tile0 = int32_t(nt);
      p.tiles_n = tn;
      p.tiles_mn = tm * tn;
      nt += int64_t(tm) * tn * p.ksplit;
    }
    TTR_REQUIRE(nt < (int64_t(1) << 31), ST_INVALID_ARG, "GEMM batch too large");
    ntiles[ci] = nt;
    std::memcpy(pack.data() + seg_bytes + ids_bytes + prob_bytes * ci, probs.data(), probs.size() * sizeof(GemmProb));
  }
  const bool inl = !tune && ctx.args_ok() &&
                   probs.size() <= size_t(GEMM_INL_PROBS) && segs.size() <= size_t(GEMM_INL_SEGS);
  const char* dpack = inl ? nullptr : ctx.upload(pack);
  const GemmSeg* ds = reinterpret_cast<const GemmSeg*>(dpack);
  const int32_t* di = reinterpret_cast<const int32_t*>(dpack + seg_bytes);
  auto dprobs = [&](size_t ci) { return reinterpret_cast<const GemmProb*>(dpack + seg_bytes + ids_bytes + prob_bytes * ci); };
  auto reduce = [&](const GemmProb* dp, const int32_t* dids, const auto& a) {
    if (split_ids.empty()) return;
    constexpr bool I = std::is_same<std::decay_t<decltype(a)>, GemmInl<true>>::value;
    int64_t mx = 1;
    for (int32_t i : split_ids) mx = std::max<int64_t>(mx, int64_t(probs[size_t(i)].M) * probs[size_t(i)].N);
    dim3 grid(unsigned(std::min<int64_t>(256, (mx + 255) / 256)), unsigned(split_ids.size())), block(256);
    if (tc == Prec::F64) TTR_LAUNCH(ctx, (gemm_splitk_reduce<double, I>), grid, block, 0, dp, dids, a);
    else if (tc == Prec::F32) TTR_LAUNCH(ctx, (gemm_splitk_reduce<float, I>), grid, block, 0, dp, dids, a);
    else TTR_LAUNCH(ctx, (gemm_splitk_reduce<bf16s, I>), grid, block, 0, dp, dids, a);
    TTR_HIP(hipGetLastError());
  };
  static const char* names[3] = {"gemm_f64", "gemm_f32", "gemm_bf16"};
  auto log_launch = [&](int c, int64_t nt) {
    if (ctx.launch_log)
      std::fprintf(ctx.launch_log, "%s %s cfg=%d wgs=%lld probs=%d flops=%.0f bytes=%.0f\n", site ? site : "-",
                   names[int(math)], c, (long long)nt, np, flops, bytes);
  };
  const std::string site_name = site ? std::string("site:") + site : std::string();
  ProfScope ps(ctx, names[int(math)], flops, bytes, site ? site_name.c_str() : nullptr);
  if (tune) {
    hipEvent_t e0, e1;
    TTR_HIP(hipEventCreate(&e0));
    TTR_HIP(hipEventCreate(&e1));
    float best = 0.0f;
    int best_cfg = cfg;
    for (size_t ci = 0; ci < cands.size(); ++ci) {
      TTR_HIP(hipEventRecord(e0, ctx.stream));
      log_launch(cands[ci], ntiles[ci]);
      launch(ctx, ta, tb, tc, math, cands[ci], int(ntiles[ci]), dprobs(ci), np, ds, GemmInl<false>{});
      TTR_HIP(hipEventRecord(e1, ctx.stream));
      TTR_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      TTR_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (best == 0.0f || ms < best) {
        best = ms;
        best_cfg = cands[ci];
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx.gemm_cfg[key] = best_cfg;  // every candidate wrote the same C (or the same partial sums)
    reduce(dprobs(0), di, GemmInl<false>{});
    return;
  }
  log_launch(cfg, ntiles[0]);
  if (inl) {
    GemmInl<true> d{};
    std::memcpy(d.p, probs.data(), probs.size() * sizeof(GemmProb));
    std::memcpy(d.s, segs.data(), segs.size() * sizeof(GemmSeg));
    std::copy(split_ids.begin(), split_ids.end(), d.ids);
    launch(ctx, ta, tb, tc, math, cfg, int(ntiles[0]), nullptr, np, nullptr, d);
    reduce(nullptr, nullptr, d);
    return;
  }
  launch(ctx, ta, tb, tc, math, cfg, int(ntiles[0]), dprobs(0), np, ds, GemmInl<false>{});
  reduce(dprobs(0), di, GemmInl<false>{});
}

}  // namespace ttr
