// Batched one-sided (Hestenes) Jacobi SVD in FP64, block variant.
//
// One workgroup per matrix B (p x p).  Columns are grouped in blocks of 32; a
// sweep visits every pair of blocks once in round-robin order.  A visit
//   1. forms the Gram matrix G = C^T C of the (up to) 64 columns C of the
//      block pair in LDS, straight from the current columns (so the rotation
//      angles always come from fresh inner products, as in plain one-sided
//      Jacobi);
//   2. if some pair is not yet orthogonal (|G_ij| > tol sqrt(G_ii G_jj),
//      tol = p eps), diagonalises G with a parallel cyclic Jacobi
//      eigen-iteration in LDS (32 disjoint rotations per round), accumulating
//      the rotation V (64 x 64);
//   3. applies V to the columns of B and of the accumulated right rotations J.
// The sweep loop ends when a whole sweep found every pair orthogonal.  On exit
// B_out = B_in J has orthogonal columns, B_out[:,c] = u_c s_c.
#include "ttr_internal.h"

namespace ttr {

namespace {

constexpr int JT = 512;         // threads per matrix
constexpr int WB = 32;          // column block width
constexpr int WC = 2 * WB;      // columns per visit
constexpr int GLD = WC + 1;     // padded LDS row
constexpr int RCH = 16;         // row chunk for Gram / apply
constexpr int MAX_SWEEPS = 40;
constexpr int MAX_INNER = 4;

__device__ __forceinline__ int rr_player(int pos, int r, int pp) {
  if (pos == 0) return 0;
  return ((pos - 1 + r) % (pp - 1)) + 1;
}

struct VisitCols {
  int col[WC];
};

__global__ __launch_bounds__(JT) void jacobi_block_kernel(const SVDJob* __restrict__ jobs) {
  const SVDJob S = jobs[blockIdx.x];
  const int p = S.p;
  const int tid = threadIdx.x;
  __shared__ double G[WC][GLD];
  __shared__ double V[WC][GLD];
  __shared__ double chunk[RCH][GLD];
  __shared__ int cols[WC];        // global column of local index, -1 if padding
  __shared__ int partner[WC];
  __shared__ double cself[WC], cpart[WC];
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
        // ---- local column map
        if (tid < WC) {
          const int blk = tid < WB ? ba : bb;
          const int c = blk < 0 ? -1 : blk * WB + (tid % WB);
          cols[tid] = (c >= 0 && c < p) ? c : -1;
        }
        __syncthreads();
        // ---- Gram matrix over row chunks
        double gacc[WC * WC / JT];
#pragma unroll
        for (int e = 0; e < WC * WC / JT; ++e) gacc[e] = 0.0;
        for (int r0 = 0; r0 < p; r0 += RCH) {
          for (int e = tid; e < RCH * WC; e += JT) {
            const int rr = e % RCH, c = e / RCH;
            const int gc = cols[c];
            chunk[rr][c] = (gc >= 0 && r0 + rr < p) ? S.B[(r0 + rr) + int64_t(gc) * S.ldb] : 0.0;
          }
          __syncthreads();
#pragma unroll
          for (int e = 0; e < WC * WC / JT; ++e) {
            const int idx = tid + e * JT;
            const int i = idx % WC, j = idx / WC;
            double s = 0.0;
#pragma unroll
            for (int rr = 0; rr < RCH; ++rr) s += chunk[rr][i] * chunk[rr][j];
            gacc[e] += s;
          }
          __syncthreads();
        }
#pragma unroll
        for (int e = 0; e < WC * WC / JT; ++e) {
          const int idx = tid + e * JT;
          const int i = idx % WC, j = idx / WC;
          G[i][j] = gacc[e];
          V[i][j] = (i == j) ? 1.0 : 0.0;
        }
        if (tid == 0) visit_rot = 0;
        __syncthreads();
        // ---- inner parallel Jacobi on G (64 indices, 63 rounds of 32 disjoint pairs)
        for (int inner = 0; inner < MAX_INNER; ++inner) {
          if (tid == 0) any_rot = 0;
          __syncthreads();
          for (int ir = 0; ir < WC - 1; ++ir) {
            if (tid == 0) round_rot = 0;
            __syncthreads();
            if (tid < WC / 2) {
              int a = rr_player(tid, ir, WC), b = rr_player(WC - 1 - tid, ir, WC);
              if (a > b) { const int t = a; a = b; b = t; }
              bool rot = false;
              double c = 1.0, s = 0.0;
              if (cols[a] >= 0 && cols[b] >= 0) {
                const double aa = G[a][a], bb2 = G[b][b], ab = G[a][b];
                if (aa > 0.0 && bb2 > 0.0 && fabs(ab) > tol * sqrt(aa) * sqrt(bb2)) {
                  const double zeta = (bb2 - aa) / (2.0 * ab);
                  const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                  c = 1.0 / sqrt(1.0 + t * t);
                  s = c * t;
                  rot = true;
                }
              }
              // column a of the rotation: c e_a - s e_b; column b: s e_a + c e_b
              partner[a] = rot ? b : -1;
              partner[b] = rot ? a : -1;
              cself[a] = c;
              cpart[a] = -s;
              cself[b] = c;
              cpart[b] = s;
              if (rot) { any_rot = 1; round_rot = 1; }
            }
            __syncthreads();
            if (round_rot) {
              double gn[WC * WC / JT], vn[WC * WC / JT];
#pragma unroll
              for (int e = 0; e < WC * WC / JT; ++e) {
                const int idx = tid + e * JT;
                const int x = idx % WC, y = idx / WC;
                const int px = partner[x], py = partner[y];
                const double cx = cself[x], cy = cself[y];
                double g = cx * cy * G[x][y];
                if (py >= 0) g += cx * cpart[y] * G[x][py];
                if (px >= 0) g += cpart[x] * cy * G[px][y];
                if (px >= 0 && py >= 0) g += cpart[x] * cpart[y] * G[px][py];
                gn[e] = g;
                double v = cy * V[x][y];
                if (py >= 0) v += cpart[y] * V[x][py];
                vn[e] = v;
              }
              __syncthreads();
#pragma unroll
              for (int e = 0; e < WC * WC / JT; ++e) {
                const int idx = tid + e * JT;
                const int x = idx % WC, y = idx / WC;
                G[x][y] = gn[e];
                V[x][y] = vn[e];
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
        // ---- apply V to the visited columns of B and J: C <- C V
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
            for (int e = tid; e < RCH * WC; e += JT) {
              const int rr = e % RCH, c = e / RCH;
              const int gc = cols[c];
              if (gc < 0 || r0 + rr >= p) continue;
              double s = 0.0;
#pragma unroll 16
              for (int u = 0; u < WC; ++u) s += chunk[rr][u] * V[u][c];
              M[(r0 + rr) + int64_t(gc) * ld] = s;
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
  const int lane = tid & 63, wave = tid >> 6;
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
    fl += 9.0 * double(j.p) * j.p * j.p;  // one sweep of rotations (nominal)
    by += 16.0 * double(j.p) * j.p;
  }
  ProfScope ps(ctx, "jacobi_svd", fl, by);
  // (sweep counts are read back by the caller through SVDJob::info)
  hipLaunchKernelGGL(jacobi_block_kernel, dim3(unsigned(live.size())), dim3(JT), 0, ctx.stream, d);
  TTR_HIP(hipGetLastError());
}

}  // namespace ttr
