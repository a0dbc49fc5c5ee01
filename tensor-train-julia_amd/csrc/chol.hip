// Batched Cholesky certificate that a truncated SVD keeps every singular value.
//
// The SVD sweep of tt_rounding keeps, at each bond, the largest k such that the
// dropped singular values carry less than tol^2 of the squared Frobenius norm
// (floor, src/tt_rounding.jl:297-315).  It drops nothing exactly when
//   s_min^2 >= tol^2 * sum_i s_i^2,
// i.e. when G - tau I is positive definite for G = B^T B and
// tau = tol^2 trace(G).  This kernel attempts the Cholesky factorisation of
// G - tau' I with tau' = tau (1 + 2^-6) + 8 (m + 1) u trace(G) (u = 2^-53): the
// last term bounds the rounding error of forming G in FP64 and of the
// factorisation itself (Demmel's bound |E| <= gamma_{m+1} |L||L^T|, whose norm
// is at most trace(G) (1 + O(mu))), so success proves that the exact singular
// values of B all lie above the floor rule's threshold with a margin that no
// correct SVD can cross.  Then the bond keeps rank m and the SVD can be skipped:
// any orthonormal basis of the row (column) space gives the same represented
// tensor, so the QR/LQ factor already computed serves as the orthogonal core.
// A failed certificate decides nothing; the caller then runs the SVD.
//
// One workgroup per matrix (256 threads), right-looking with panels of 16
// columns: wave 0 factors the 16 x 16 diagonal block in registers (cross-lane
// moves, no barriers), every thread solves rows of the panel below against it
// into LDS (transposed, Pt[c][i]), and the lower triangle of the trailing
// matrix is updated in global memory in 4 x 4 tiles, FP64 FMA (on gfx950 the
// FP64 VALU rate equals the FP64 MFMA rate, and the panel width keeps the
// operands in LDS).  The factor itself is not kept.
#include "ttr_internal.h"

namespace ttr {

namespace {

constexpr int CT = 256;             // threads per matrix
constexpr int CNB = 16;             // panel width
constexpr int CPAD = CHOL_MAX_M + 8;  // LDS panel rows (tiles read up to 3 rows past m)

__device__ __forceinline__ double shfl_d(double v, int src) {
  const int lo = __shfl(__double2loint(v), src);
  const int hi = __shfl(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

__global__ __launch_bounds__(CT) void chol_certify_kernel(const CholJob* __restrict__ jobs) {
  const CholJob J = jobs[blockIdx.x];
  __shared__ double Pt[CNB][CPAD];    // the current panel below the diagonal block, Pt[c][i]
  __shared__ double L11[CNB][CNB + 1];
  __shared__ double rdiag[CNB];
  __shared__ double red[CT / 64];
  __shared__ int fail;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m = J.m;
  double* A = J.G;
  const int64_t ld = J.ld;

  double tr = 0.0;
  for (int i = tid; i < m; i += CT) tr += A[i + int64_t(i) * ld];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) tr += __shfl_xor(tr, off);
  if (lane == 0) red[wave] = tr;
  __syncthreads();
  tr = red[0] + red[1] + red[2] + red[3];
  const double u = 0x1p-53;
  const double tau = J.tol * J.tol * tr * (1.0 + 0x1p-6) + 8.0 * double(m + 1) * u * tr;
  if (!(tr > 0.0) || !isfinite(tr) || !isfinite(tau) || m > CHOL_MAX_M) {
    if (tid == 0) *J.flag = 0;
    return;
  }

  for (int k0 = 0; k0 < m; k0 += CNB) {
    const int nb = min(CNB, m - k0);
    const int k1 = k0 + nb;
    // ---- diagonal block D - tau I (padded with the identity), factored by wave 0.
    // Lane l holds row r = l & 15 of columns 4g .. 4g + 3, g = l >> 4.
    if (wave == 0) {
      const int r = lane & 15, g = lane >> 4;
      double dv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = 4 * g + q;
        if (r < nb && j < nb) dv[q] = A[(k0 + r) + int64_t(k0 + j) * ld] - (r == j ? tau : 0.0);
        else dv[q] = r == j ? 1.0 : 0.0;
      }
      bool bad = false;
#pragma unroll
      for (int c = 0; c < CNB; ++c) {
        const int cg = c >> 2, cq = c & 3;
        const double piv = shfl_d(dv[cq], c + 16 * cg);  // D(c, c): same value on every lane
        if (!(piv > 0.0) || !isfinite(piv)) { bad = true; break; }
        const double s = sqrt(piv), rs = 1.0 / s;
        if (g == cg) {
          if (r > c) dv[cq] *= rs;
          else if (r == c) dv[cq] = s;
        }
        const double lr = shfl_d(dv[cq], r + 16 * cg);  // L(r, c)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int j = 4 * g + q;
          const double lj = shfl_d(dv[cq], j + 16 * cg);  // L(j, c)
          if (j > c && r >= j) dv[q] -= lr * lj;
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) L11[r][4 * g + q] = dv[q];
      if (lane < CNB) rdiag[lane] = 0.0;  // set below from the diagonal
      if (lane == 0) fail = bad ? 1 : 0;
    }
    __syncthreads();
    if (fail) {
      if (tid == 0) *J.flag = 0;
      return;
    }
    if (k1 >= m) break;
    if (tid < CNB) rdiag[tid] = 1.0 / L11[tid][tid];
    __syncthreads();
    // ---- panel rows below: L21 = A21 L11^-T, into Pt (rows m .. m + 7 zeroed)
    for (int i = k1 + tid; i < m + 8; i += CT) {
      double a[CNB];
#pragma unroll
      for (int c = 0; c < CNB; ++c) a[c] = (i < m && c < nb) ? A[i + int64_t(k0 + c) * ld] : 0.0;
#pragma unroll
      for (int c = 0; c < CNB; ++c) {
        double s = a[c];
#pragma unroll
        for (int q = 0; q < c; ++q) s -= a[q] * L11[c][q];
        a[c] = s * rdiag[c];
      }
#pragma unroll
      for (int c = 0; c < CNB; ++c) Pt[c][i] = a[c];
    }
    __syncthreads();
    // ---- trailing update of the lower triangle of rows / columns [k1, m):
    // 4 x 4 tiles listed column by column (tile column tj holds the row tiles
    // ti = tj .. nt-1), so consecutive threads touch consecutive rows of one column
    const int nt = (m - k1 + 3) / 4;
    const int64_t ntiles = int64_t(nt) * (nt + 1) / 2;
    const double b2 = 2.0 * nt + 1.0;
    for (int64_t e = tid; e < ntiles; e += CT) {
      // column tj: the largest tj with start(tj) = tj nt - tj (tj - 1) / 2 <= e
      int tj = int((b2 - sqrt(b2 * b2 - 8.0 * double(e))) * 0.5);
      tj = max(0, min(tj, nt - 1));
      while (tj > 0 && int64_t(tj) * nt - int64_t(tj) * (tj - 1) / 2 > e) --tj;
      while (tj + 1 < nt && int64_t(tj + 1) * nt - int64_t(tj + 1) * tj / 2 <= e) ++tj;
      const int ti = tj + int(e - (int64_t(tj) * nt - int64_t(tj) * (tj - 1) / 2));
      const int i0 = k1 + 4 * ti, j0 = k1 + 4 * tj;
      double acc[4][4];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
      for (int c = 0; c < nb; ++c) {
        double pi[4], pj[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          pi[a] = Pt[c][i0 + a];
          pj[a] = Pt[c][j0 + a];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) acc[a][b] = fma(pi[a], pj[b], acc[a][b]);
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int j = j0 + b;
        if (j >= m) continue;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int i = i0 + a;
          if (i < m && i >= j) A[i + int64_t(j) * ld] -= acc[a][b];
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0) *J.flag = 1;
}

}  // namespace

void chol_certify_batched(Ctx& ctx, const std::vector<CholJob>& jobs) {
  if (jobs.empty()) return;
  for (const auto& j : jobs)
    TTR_REQUIRE(j.m >= 1 && j.m <= CHOL_MAX_M && j.ld >= j.m, ST_INVALID_ARG, "chol_certify: bad shape");
  const CholJob* d = ctx.upload(jobs);
  double fl = 0.0;
  for (const auto& j : jobs) fl += double(j.m) * j.m * j.m / 3.0;
  ProfScope ps(ctx, "svd_certify", fl, 0.0);
  TTR_LAUNCH(ctx, chol_certify_kernel, dim3(unsigned(jobs.size())), dim3(CT), 0, d);
  TTR_HIP(hipGetLastError());
}

}  // namespace ttr
