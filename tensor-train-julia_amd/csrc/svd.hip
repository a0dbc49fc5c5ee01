// Batched one-sided (Hestenes) Jacobi SVD in FP64, block variant.
//
// One workgroup per matrix B (p x p).  Columns are grouped in blocks of 32; a
// sweep visits every pair of blocks once in round-robin order.  A visit
//   1. forms the Gram matrix G = C^T C of the (up to) 64 columns C of the
//      block pair in LDS with FP64 MFMA, straight from the current columns
//      (so rotation angles always come from fresh inner products, as in plain
//      one-sided Jacobi);
//   2. if some pair is not yet orthogonal (|G_ij| > tol sqrt(G_ii G_jj),
//      tol = p eps), runs a parallel cyclic Jacobi eigen-iteration on G in LDS
//      (32 disjoint rotations per round, each 2x2 block of G and each row pair
//      of V updated in place by one thread), accumulating the rotation V;
//   3. applies V to the visited columns of B and of the accumulated right
//      rotations J with FP64 MFMA (C <- C V, 32-row chunks staged in LDS).
// The sweep loop ends when a whole sweep found every pair orthogonal.  On exit
// B_out = B_in J has orthogonal columns, B_out[:,c] = u_c s_c.
#include "ttr_internal.h"

namespace ttr {

namespace {

typedef double d4v __attribute__((ext_vector_type(4)));

constexpr int JT = 512;         // threads per matrix (8 waves)
constexpr int WB = 32;          // column block width
constexpr int WC = 2 * WB;      // columns per visit
constexpr int GLD = WC + 1;     // padded LDS row
constexpr int RCH = 32;         // row chunk for Gram / apply
constexpr int MAX_SWEEPS = 40;
constexpr int MAX_INNER = 2;
constexpr int NPAIR = WC / 2;

__device__ __forceinline__ int rr_player(int pos, int r, int pp) {
  if (pos == 0) return 0;
  return ((pos - 1 + r) % (pp - 1)) + 1;
}

__global__ __launch_bounds__(JT) void jacobi_block_kernel(const SVDJob* __restrict__ jobs) {
  const SVDJob S = jobs[blockIdx.x];
  const int p = S.p;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ double G[WC][GLD];
  __shared__ double V[WC][GLD];
  __shared__ double chunk[RCH][GLD];
  __shared__ int cols[WC];  // global column of local index, -1 if padding
  __shared__ int pa[NPAIR], pb[NPAIR];
  __shared__ double pc[NPAIR], ps[NPAIR];
  __shared__ int any_rot, visit_rot, sweep_rot, round_rot;

  for (int64_t idx = tid; idx < int64_t(p) * p; idx += JT) {
    const int64_t r = idx % p, c = idx / p;
    S.J[r + c * S.ldj] = (r == c) ? 1.0 : 0.0;
  }
  // inner products of length p carry rounding noise of order sqrt(p)..p eps, so a
  // pair counts as orthogonal below p eps (at least 16 eps)
  const double tol = double(p > 16 ? p : 16) * 2.220446049250313e-16;
  int sweeps_done = 0;
  const int nblk = (p + WB - 1) / WB;
  const int pp = nblk + (nblk & 1);
  __syncthreads();

  for (int sweep = 0; sweep < MAX_SWEEPS && p > 1; ++sweep) {
    sweeps_done = sweep + 1;
    if (tid == 0) sweep_rot = 0;
    __syncthreads();
    const int rounds = (nblk == 1) ? 1 : pp - 1;
    for (int r = 0; r < rounds; ++r) {
      for (int k = 0; k < ((nblk == 1) ? 1 : pp / 2); ++k) {
        int ba, bb;
        if (nblk == 1) { ba = 0; bb = -1; }
        else {
          ba = rr_player(k, r, pp);
          bb = rr_player(pp - 1 - k, r, pp);
          if (ba >= nblk) ba = -1;
          if (bb >= nblk) bb = -1;
          if (ba < 0 && bb < 0) continue;
        }
        if (tid < WC) {
          const int blk = tid < WB ? ba : bb;
          const int c = blk < 0 ? -1 : blk * WB + (tid % WB);
          cols[tid] = (c >= 0 && c < p) ? c : -1;
        }
        __syncthreads();
        // ---- Gram matrix G = C^T C (MFMA f64 16x16x4; wave w owns output blocks w and w+8)
        d4v gacc[2] = {d4v{0, 0, 0, 0}, d4v{0, 0, 0, 0}};
        for (int r0 = 0; r0 < p; r0 += RCH) {
          for (int e = tid; e < RCH * WC; e += JT) {
            const int rr = e % RCH, c = e / RCH;
            const int gc = cols[c];
            chunk[rr][c] = (gc >= 0 && r0 + rr < p) ? S.B[(r0 + rr) + int64_t(gc) * S.ldb] : 0.0;
          }
          __syncthreads();
#pragma unroll
          for (int t = 0; t < 2; ++t) {
            const int blk = wave + 8 * t, bi = blk >> 2, bj = blk & 3;
#pragma unroll
            for (int kk = 0; kk < RCH; kk += 4) {
              const double a = chunk[kk + (lane >> 4)][16 * bi + (lane & 15)];
              const double b = chunk[kk + (lane >> 4)][16 * bj + (lane & 15)];
              gacc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, gacc[t], 0, 0, 0);
            }
          }
          __syncthreads();
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const int blk = wave + 8 * t, bi = blk >> 2, bj = blk & 3;
#pragma unroll
          for (int q = 0; q < 4; ++q) G[16 * bi + 4 * q + (lane >> 4)][16 * bj + (lane & 15)] = gacc[t][q];
        }
        for (int e = tid; e < WC * WC; e += JT) V[e % WC][e / WC] = (e % WC == e / WC) ? 1.0 : 0.0;
        if (tid == 0) visit_rot = 0;
        __syncthreads();
        // ---- inner parallel Jacobi on G
        for (int inner = 0; inner < MAX_INNER; ++inner) {
          if (tid == 0) any_rot = 0;
          for (int ir = 0; ir < WC - 1; ++ir) {
            if (tid == 0) round_rot = 0;
            __syncthreads();
            if (tid < NPAIR) {
              int a = rr_player(tid, ir, WC), b = rr_player(WC - 1 - tid, ir, WC);
              if (a > b) { const int t = a; a = b; b = t; }
              double c = 1.0, s = 0.0;
              if (cols[a] >= 0 && cols[b] >= 0) {
                const double aa = G[a][a], bb2 = G[b][b], ab = G[a][b];
                if (aa > 0.0 && bb2 > 0.0 && fabs(ab) > tol * sqrt(aa) * sqrt(bb2)) {
                  const double zeta = (bb2 - aa) / (2.0 * ab);
                  const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                  c = 1.0 / sqrt(1.0 + t * t);
                  s = c * t;
                  any_rot = 1;
                  round_rot = 1;
                }
              }
              pa[tid] = a;
              pb[tid] = b;
              pc[tid] = c;
              ps[tid] = s;
            }
            __syncthreads();
            if (round_rot) {
              // G <- R^T G R with R the block diagonal of [[c, s], [-s, c]] on (a_k, b_k);
              // each 2x2 block (pair kr rows, pair kc columns) is owned by one thread
              for (int e = tid; e < NPAIR * NPAIR; e += JT) {
                const int kr = e % NPAIR, kc = e / NPAIR;
                const int ar = pa[kr], br = pb[kr], ac = pa[kc], bc = pb[kc];
                const double cr = pc[kr], sr = ps[kr], cc = pc[kc], sc = ps[kc];
                const double g00 = G[ar][ac], g01 = G[ar][bc], g10 = G[br][ac], g11 = G[br][bc];
                const double t00 = cr * g00 - sr * g10, t01 = cr * g01 - sr * g11;
                const double t10 = sr * g00 + cr * g10, t11 = sr * g01 + cr * g11;
                G[ar][ac] = t00 * cc - t01 * sc;
                G[ar][bc] = t00 * sc + t01 * cc;
                G[br][ac] = t10 * cc - t11 * sc;
                G[br][bc] = t10 * sc + t11 * cc;
              }
              // V <- V R: row x, pair l
              for (int e = tid; e < WC * NPAIR; e += JT) {
                const int x = e % WC, l = e / WC;
                const int a = pa[l], b = pb[l];
                const double c = pc[l], s = ps[l];
                const double va = V[x][a], vb = V[x][b];
                V[x][a] = c * va - s * vb;
                V[x][b] = s * va + c * vb;
              }
            }
            __syncthreads();
          }
          const int ar = any_rot;
          if (tid == 0 && ar) visit_rot = 1;
          __syncthreads();
          if (!ar) break;
        }
        const int vr = visit_rot;
        __syncthreads();
        if (!vr) continue;
        if (tid == 0) sweep_rot = 1;
        // ---- apply V to the visited columns of B and J: C <- C V (wave w: row block w>>2, column block w&3)
        for (int which = 0; which < 2; ++which) {
          double* M = which == 0 ? S.B : S.J;
          const int64_t ld = which == 0 ? S.ldb : S.ldj;
          for (int r0 = 0; r0 < p; r0 += RCH) {
            for (int e = tid; e < RCH * WC; e += JT) {
              const int rr = e % RCH, c = e / RCH;
              const int gc = cols[c];
              chunk[rr][c] = (gc >= 0 && r0 + rr < p) ? M[(r0 + rr) + int64_t(gc) * ld] : 0.0;
            }
            __syncthreads();
            const int bi = wave >> 2, bj = wave & 3;
            d4v acc = {0, 0, 0, 0};
#pragma unroll
            for (int kk = 0; kk < WC; kk += 4) {
              const double a = chunk[16 * bi + (lane & 15)][kk + (lane >> 4)];
              const double b = V[kk + (lane >> 4)][16 * bj + (lane & 15)];
              acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int row = r0 + 16 * bi + 4 * q + (lane >> 4);
              const int gc = cols[16 * bj + (lane & 15)];
              if (gc >= 0 && row < p) M[row + int64_t(gc) * ld] = acc[q];
            }
            __syncthreads();
          }
        }
      }
    }
    const int sr = sweep_rot;
    __syncthreads();
    if (!sr) break;
  }
  // singular values: column norms
  for (int c = wave; c < p; c += JT / 64) {
    const double* x = S.B + int64_t(c) * S.ldb;
    double s = 0.0;
    for (int i = lane; i < p; i += 64) s += x[i] * x[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) S.sv[c] = sqrt(s);
  }
  if (tid == 0 && S.info) S.info[0] = sweeps_done;
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
    fl += 9.0 * double(j.p) * j.p * j.p;  // one sweep of plain one-sided Jacobi (nominal)
    by += 16.0 * double(j.p) * j.p;
  }
  ProfScope ps(ctx, "jacobi_svd", fl, by);
  hipLaunchKernelGGL(jacobi_block_kernel, dim3(unsigned(live.size())), dim3(JT), 0, ctx.stream, d);
  TTR_HIP(hipGetLastError());
}

}  // namespace ttr
