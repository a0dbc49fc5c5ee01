// Batched one-sided (Hestenes) Jacobi SVD in FP64.
//
// One workgroup (16 wavefronts) per matrix B (p x p).  A sweep visits every
// column pair once in round-robin order (p-1 rounds of p/2 disjoint pairs);
// each wavefront rotates its pairs with wave-shuffle reductions for the three
// inner products, then the workgroup synchronises before the next round.  The
// rotations are accumulated in J, so B_in = (B_out) J^T with orthogonal
// columns in B_out: B_out[:,c] = u_c s_c.  Convergence: no pair with
// |b_i.b_j| > tol * |b_i| |b_j| in a whole sweep, tol = sqrt(p) * eps.
#include "ttr_internal.h"

namespace ttr {

namespace {

constexpr int JT = 1024;
constexpr int JW = JT / 64;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// player at position pos in round r of the circle method with pp (even) players
__device__ __forceinline__ int rr_player(int pos, int r, int pp) {
  if (pos == 0) return 0;
  return ((pos - 1 + r) % (pp - 1)) + 1;
}

__global__ __launch_bounds__(JT) void jacobi_kernel(const SVDJob* __restrict__ jobs, int max_sweeps) {
  const SVDJob S = jobs[blockIdx.x];
  const int p = S.p;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ int rotated;
  for (int64_t idx = tid; idx < int64_t(p) * p; idx += JT) {
    const int64_t r = idx % p, c = idx / p;
    S.J[r + c * S.ldj] = (r == c) ? 1.0 : 0.0;
  }
  const double tol = sqrt(double(p > 1 ? p : 1)) * 2.220446049250313e-16;
  const int pp = p + (p & 1);
  __syncthreads();
  if (p > 1) {
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
      if (tid == 0) rotated = 0;
      __syncthreads();
      for (int r = 0; r < pp - 1; ++r) {
        for (int k = wave; k < pp / 2; k += JW) {
          int a = rr_player(k, r, pp), b = rr_player(pp - 1 - k, r, pp);
          if (a >= p || b >= p) continue;
          if (a > b) { const int t = a; a = b; b = t; }
          double* x = S.B + int64_t(a) * S.ldb;
          double* y = S.B + int64_t(b) * S.ldb;
          double aa = 0.0, bb = 0.0, ab = 0.0;
          for (int i = lane; i < p; i += 64) {
            const double xi = x[i], yi = y[i];
            aa += xi * xi;
            bb += yi * yi;
            ab += xi * yi;
          }
          aa = wsum(aa);
          bb = wsum(bb);
          ab = wsum(ab);
          if (aa > 0.0 && bb > 0.0 && fabs(ab) > tol * sqrt(aa) * sqrt(bb)) {
            const double zeta = (bb - aa) / (2.0 * ab);
            const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            const double c = 1.0 / sqrt(1.0 + t * t);
            const double s = c * t;
            for (int i = lane; i < p; i += 64) {
              const double xi = x[i], yi = y[i];
              x[i] = c * xi - s * yi;
              y[i] = s * xi + c * yi;
            }
            double* u = S.J + int64_t(a) * S.ldj;
            double* v = S.J + int64_t(b) * S.ldj;
            for (int i = lane; i < p; i += 64) {
              const double ui = u[i], vi = v[i];
              u[i] = c * ui - s * vi;
              v[i] = s * ui + c * vi;
            }
            if (lane == 0) rotated = 1;
          }
        }
        __syncthreads();
      }
      const int any = rotated;
      __syncthreads();
      if (!any) break;
    }
  }
  for (int c = wave; c < p; c += JW) {
    const double* x = S.B + int64_t(c) * S.ldb;
    double s = 0.0;
    for (int i = lane; i < p; i += 64) s += x[i] * x[i];
    s = wsum(s);
    if (lane == 0) S.sv[c] = sqrt(s);
  }
}

}  // namespace

void jacobi_svd_batched(Ctx& ctx, std::vector<SVDJob>& jobs) {
  std::vector<SVDJob> live;
  for (auto& j : jobs)
    if (j.p > 0) live.push_back(j);
  if (live.empty()) return;
  const SVDJob* d = ctx.upload(live);
  double fl = 0.0, by = 0.0;
  for (const auto& j : live) {
    fl += 9.0 * double(j.p) * j.p * j.p;  // one sweep of rotations (nominal)
    by += 16.0 * double(j.p) * j.p;
  }
  ProfScope ps(ctx, "jacobi_svd", fl, by);
  hipLaunchKernelGGL(jacobi_kernel, dim3(unsigned(live.size())), dim3(JT), 0, ctx.stream, d, 60);
  TTR_HIP(hipGetLastError());
}

}  // namespace ttr
