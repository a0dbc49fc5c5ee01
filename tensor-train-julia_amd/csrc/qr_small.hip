// One-launch FP64 Householder QR for matrices of at most 1024 rows: one
// workgroup per matrix does the whole factorisation and the thin Q.
//
// The blocked path of qr.hip spreads every trailing update over the chip, which
// pays off for large batches of large matrices but leaves a few matrices (a
// single TT's sweep, the small bonds near the ends of a TT) bound by ~30 launch
// latencies per QR.  Here the matrix stays in L2 and one 512-thread workgroup
// (two waves per SIMD) runs, per panel of 16 columns:
//   * the panel factorisation in registers of waves 0-3 (thread t holds rows
//     t, t+256, ..; waves 4-7 wait at the barriers),
//     one workgroup reduction per column: the 16 sums x^T A(:, c) of the column
//     being reduced (x) against every panel column are formed in one pass and
//     reduced with a butterfly reduce-scatter (17 shuffles per wave instead of
//     16 full reductions), the reflector norm being the sum against itself;
//   * V (unit lower, 16 columns) written to LDS and T (dlarft) built in LDS;
//   * the trailing update A2 -= V T^T V^T A2 with v_mfma_f64_16x16x4 in slabs
//     of 16 columns per wave, all 8 waves (W = V^T A2 with two accumulator
//     chains, T^T W and the rank-16 update tile by tile);
// and then forms Q = H_1 .. H_k [I; 0] panel by panel from the last one, with
// the saved T's.  Conventions are LAPACK's (dgeqrf/dlarfg/dlarft/dorgqr):
//   beta = -sign(alpha) ||(alpha, x)||, tau = (beta - alpha)/beta,
//   v = (1, x/(alpha - beta)), tau = 0 when x == 0,
// so Q and R agree with the reference's LAPACK factorisations column by column.
#include "ttr_internal.h"
#include "qr_step.h"

namespace ttr {

namespace {

constexpr int SQT = 512;       // threads: 2 waves per SIMD for the MFMA updates
constexpr int SQW = SQT / 64;  // waves
constexpr int SPT = 256;       // threads holding the panel (waves 0-3)
constexpr int SPW = SPT / 64;
using qrstep::SNB;
using qrstep::LDV;

typedef double d4_t __attribute__((ext_vector_type(4)));

struct SmallQRJob {
  double* A;
  int64_t lda;
  int32_t m, n;
  double* Q;      // thin Q (m x min(m, n)), or null
  int64_t ldq;
  double* tau;    // min(m, n)
  double* T;      // one 16 x 16 T per panel (column-major), read back when forming Q
};

// C(rows 0..m2-1, 16 columns j0..) -= V (T^T or T) V^T C for the 16-column
// slab starting at j0 (columns >= ncol are left alone), C column-major with ld
// ldc.  trans: W2 = T^T W (factorisation), else W2 = T W (forming Q).
// Loads are unconditional global-address-space loads at clamped (always valid)
// row indices, so the next batch is in flight while the current one is
// multiplied: rows past m2 meet zero rows of V (W) or are never stored (update).
// Lanes of columns past ncol read column j0 and store nothing.
typedef const double __attribute__((address_space(1)))* gcd_t;
typedef double __attribute__((address_space(1)))* gd_t;

__device__ __forceinline__ void apply_panel_slab(const double* __restrict__ Vs, const double* __restrict__ Ts,
                                                 double* C, int64_t ldc, int m2, int j0, int ncol, bool trans,
                                                 int lane) {
  const int li = lane & 15, lk = lane >> 4;
  const int j = j0 + li;
  const bool colok = j < ncol;
  gd_t Cg = (gd_t)(C + ldc * int64_t(colok ? j : j0));
  // W = V^T C (16 x 16): batches of 8 k-steps (32 rows), two accumulator chains.
  // Vs rows past m2 are zero up to a multiple of 64 rows (the loop may run one
  // batch past m2 and reads Vs rows below roundup(m2, 64) <= 256 R).
  constexpr int U = 8;
  d4_t w0 = {0.0, 0.0, 0.0, 0.0}, w1 = {0.0, 0.0, 0.0, 0.0};
  auto load_b = [&](double (&b)[U], int kb) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      b[u] = Cg[min(kb + 4 * u + lk, m2 - 1)];
    }
  };
  auto mma_b = [&](const double (&b)[U], int kb) {
    double a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = Vs[(kb + 4 * u + lk) * LDV + li];
#pragma unroll
    for (int u = 0; u < U; u += 2) {
      w0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], w0, 0, 0, 0);
      w1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u + 1], b[u + 1], w1, 0, 0, 0);
    }
  };
  {
    double ba[U], bb[U];
    load_b(ba, 0);
    for (int kb = 0; kb < m2; kb += 8 * U) {
      load_b(bb, kb + 4 * U);
      mma_b(ba, kb);
      load_b(ba, kb + 8 * U);
      mma_b(bb, kb + 4 * U);
    }
  }
  const d4_t w = w0 + w1;
  // W2 = T^T W or T W: B operand rows 4r..4r+3 are register r of w
  d4_t w2 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int k = 4 * r + lk;
    const double a = trans ? Ts[k * LDV + li] : Ts[li * LDV + k];
    w2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, w[r], w2, 0, 0, 0);
  }
  // C -= V W2 in batches of 4 tiles of 16 rows (register r of a tile is row
  // 4 r + lane / 16), the next batch loaded before the current one is updated
  auto load_c = [&](d4_t (&c)[4], int i0) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + 16 * t + 4 * r + lk;
        c[t][r] = Cg[min(i, m2 - 1)];
      }
  };
  auto update_c = [&](d4_t (&c)[4], int i0) {
    double a[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int t = 0; t < 4; ++t) a[r][t] = -Vs[(i0 + 16 * t + li) * LDV + 4 * r + lk];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int t = 0; t < 4; ++t) c[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r][t], w2[r], c[t], 0, 0, 0);
    if (colok) {
      if (i0 + 64 <= m2) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r) Cg[i0 + 16 * t + 4 * r + lk] = c[t][r];
      } else {
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = i0 + 16 * t + 4 * r + lk;
            if (i < m2) Cg[i] = c[t][r];
          }
      }
    }
  };
  {
    d4_t ca[4], cb[4];
    load_c(ca, 0);
    for (int i0 = 0; i0 < m2; i0 += 128) {
      load_c(cb, i0 + 64);
      update_c(ca, i0);
      if (i0 + 64 >= m2) break;
      load_c(ca, i0 + 128);
      update_c(cb, i0 + 64);
    }
  }
}

template <int R>
__global__ __launch_bounds__(SQT) void qr_small_kernel(const SmallQRJob* __restrict__ jobs) {
  const SmallQRJob J = jobs[blockIdx.x];
  __shared__ double Vs[SPT * R * LDV];
  __shared__ qrstep::PanelShared16<SPT> sh;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m = J.m, n = J.n, kq = min(m, n);
  double* A = J.A;
  const int64_t lda = J.lda;
  const int npanel = (kq + SNB - 1) / SNB;

  for (int p = 0; p < npanel; ++p) {
    const int c0 = p * SNB, nb = min(SNB, kq - c0), m2 = m - c0;
    double* A0 = A + c0 + lda * c0;  // the panel's top-left element
    const bool pt = tid < SPT;
    double P[R][SNB];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = tid + SPT * q;
#pragma unroll
      for (int c = 0; c < SNB; ++c) P[q][c] = (pt && i < m2 && c < nb) ? A0[i + lda * c] : 0.0;
    }
    for (int e = tid; e < SNB * LDV; e += SQT) sh.Ts[e] = 0.0;
    __syncthreads();
    qrstep::PanelStep16<R, SPT, 0>::run(P, sh, J.tau + c0, nb, m2, tid, lane, wave);
    // R and V back to A; V (unit lower, zero above, zero rows past m2) to LDS
    if (pt) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = tid + SPT * q;
#pragma unroll
        for (int c = 0; c < SNB; ++c) {
          if (i < m2 && c < nb) A0[i + lda * c] = P[q][c];
          Vs[i * LDV + c] = (i < m2 && c < nb) ? (i > c ? P[q][c] : (i == c ? 1.0 : 0.0)) : 0.0;
        }
      }
    }
    __syncthreads();
    if (J.Q && tid < SNB * SNB) J.T[size_t(p) * SNB * SNB + tid] = sh.Ts[(tid % SNB) * LDV + tid / SNB];
    // trailing columns c0+nb .. n-1 in slabs of 16, one slab per wave at a time
    const int first = c0 + nb;
    const int nslab = (n - first + SNB - 1) / SNB;
    for (int sl = wave; sl < nslab; sl += SQW)
      apply_panel_slab(Vs, sh.Ts, A + c0 + lda * first, lda, m2, sl * SNB, n - first, true, lane);
    __syncthreads();
  }
  if (!J.Q) return;
  // thin Q: identity, then H_p applied for p = last .. first
  for (int64_t e = tid; e < int64_t(m) * kq; e += SQT) {
    const int i = int(e % m), c = int(e / m);
    J.Q[i + J.ldq * c] = (i == c) ? 1.0 : 0.0;
  }
  for (int p = npanel - 1; p >= 0; --p) {
    const int c0 = p * SNB, nb = min(SNB, kq - c0), m2 = m - c0;
    const double* A0 = A + c0 + lda * c0;
    for (int e = tid; e < SPT * R * SNB; e += SQT) {
      const int i = e / SNB, c = e % SNB;
      Vs[i * LDV + c] = (i < m2 && c < nb) ? (i > c ? A0[i + lda * c] : (i == c ? 1.0 : 0.0)) : 0.0;
    }
    for (int e = tid; e < SNB * SNB; e += SQT) sh.Ts[(e % SNB) * LDV + e / SNB] = J.T[size_t(p) * SNB * SNB + e];
    __syncthreads();
    const int nslab = (kq - c0 + SNB - 1) / SNB;
    for (int sl = wave; sl < nslab; sl += SQW)
      apply_panel_slab(Vs, sh.Ts, J.Q + c0 + J.ldq * c0, J.ldq, m2, sl * SNB, kq - c0, false, lane);
    __syncthreads();
  }
}

}  // namespace

void qr_small_batched(Ctx& ctx, const std::vector<QRJob>& jobs) {
  std::vector<SmallQRJob> sj;
  int max_m = 0;
  double fl = 0.0, by = 0.0;
  for (const auto& J : jobs) {
    TTR_REQUIRE(J.m >= 0 && J.n >= 0 && J.lda >= std::max(1, J.m), ST_INVALID_ARG, "qr: bad shape");
    TTR_REQUIRE(J.m <= QR_SMALL_MAX_M, ST_INVALID_ARG, "qr_small: more than 1024 rows");
    const int kq = std::min(J.m, J.n);
    if (kq == 0) continue;
    const int npanel = (kq + SNB - 1) / SNB;
    SmallQRJob s{J.A, J.lda, J.m, J.n, J.Q, J.ldq, ctx.ws.alloc_n<double>(size_t(kq)), nullptr};
    if (J.Q) s.T = ctx.ws.alloc_n<double>(size_t(npanel) * SNB * SNB);
    sj.push_back(s);
    max_m = std::max(max_m, J.m);
    const double m = J.m, n = J.n, k = kq;
    fl += 2.0 * m * n * k - (m + n) * k * k + 2.0 * k * k * k / 3.0;
    by += 16.0 * m * n;
    if (J.Q) {
      fl += 2.0 * m * k * k - 2.0 * k * k * k / 3.0;
      by += 8.0 * m * k;
    }
  }
  if (sj.empty()) return;
  const SmallQRJob* d = ctx.upload(sj);
  const int R = (max_m + SPT - 1) / SPT;
  ProfScope ps(ctx, "qr_small", fl, by);
  switch (R) {
    case 0:
    case 1: TTR_LAUNCH(ctx, qr_small_kernel<1>, dim3(unsigned(sj.size())), dim3(SQT), 0, d); break;
    case 2: TTR_LAUNCH(ctx, qr_small_kernel<2>, dim3(unsigned(sj.size())), dim3(SQT), 0, d); break;
    case 3: TTR_LAUNCH(ctx, qr_small_kernel<3>, dim3(unsigned(sj.size())), dim3(SQT), 0, d); break;
    default: TTR_LAUNCH(ctx, qr_small_kernel<4>, dim3(unsigned(sj.size())), dim3(SQT), 0, d); break;
  }
}

}  // namespace ttr
