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
constexpr int MAX_INNER = 1;
constexpr int NPAIR = WC / 2;

__device__ __forceinline__ int rr_player(int pos, int r, int pp) {
  if (pos == 0) return 0;
  return ((pos - 1 + r) % (pp - 1)) + 1;
}

__global__ __launch_bounds__(JT) void jacobi_block_kernel(const SVDJob* __restrict__ jobs) {
  const SVDJob S = jobs[blockIdx.x];
  const int p = S.p;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ double Gbuf[2][WC][GLD];  // ping-pong copies of the Gram block
  __shared__ double V[WC][GLD];
  __shared__ double chunk[RCH][GLD];
  __shared__ int cols[WC];  // global column of local index, -1 if padding

  for (int64_t idx = tid; idx < int64_t(p) * p; idx += JT) {
    const int64_t r = idx % p, c = idx / p;
    S.J[r + c * S.ldj] = (r == c) ? 1.0 : 0.0;
  }
  // inner products of length p carry rounding noise of order sqrt(p)..p eps, so a
  // pair counts as orthogonal below p eps (at least 16 eps)
  const double tol = double(p > 16 ? p : 16) * 2.220446049250313e-16;
  int sweeps_done = 0;
  uint64_t t_gram = 0, t_inner = 0, t_apply = 0, t0 = 0;
  const int nblk = (p + WB - 1) / WB;
  const int pp = nblk + (nblk & 1);
  __syncthreads();

  for (int sweep = 0; sweep < MAX_SWEEPS && p > 1; ++sweep) {
    sweeps_done = sweep + 1;
    bool sweep_rot = false;
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
        if (tid == 0) t0 = wall_clock64();
        d4v gacc[2] = {d4v{0, 0, 0, 0}, d4v{0, 0, 0, 0}};
        constexpr int PER = RCH * WC / JT;  // chunk elements per thread
        double pre[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          const int e = tid + q * JT, rr = e % RCH, gc = cols[e / RCH];
          pre[q] = (gc >= 0 && rr < p) ? S.B[rr + int64_t(gc) * S.ldb] : 0.0;
        }
        for (int r0 = 0; r0 < p; r0 += RCH) {
#pragma unroll
          for (int q = 0; q < PER; ++q) {
            const int e = tid + q * JT;
            chunk[e % RCH][e / RCH] = pre[q];
          }
          __syncthreads();
          if (r0 + RCH < p) {  // prefetch the next chunk while the MFMAs run
#pragma unroll
            for (int q = 0; q < PER; ++q) {
              const int e = tid + q * JT, rr = r0 + RCH + e % RCH, gc = cols[e / RCH];
              pre[q] = (gc >= 0 && rr < p) ? S.B[rr + int64_t(gc) * S.ldb] : 0.0;
            }
          }
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
          for (int q = 0; q < 4; ++q) Gbuf[0][16 * bi + 4 * q + (lane >> 4)][16 * bj + (lane & 15)] = gacc[t][q];
        }
        for (int e = tid; e < WC * WC; e += JT) V[e % WC][e / WC] = (e % WC == e / WC) ? 1.0 : 0.0;
        __syncthreads();
        // scale G by a power of two (rotation angles are scale free) so that the
        // squares below cannot overflow for unnormalised inputs
        {
          double mx = 0.0;
          for (int i = lane; i < WC; i += 64) mx = fmax(mx, Gbuf[0][i][i]);
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
          const double sc = (mx > 0.0 && isfinite(mx)) ? ldexp(1.0, -ilogb(mx)) : 1.0;
          __syncthreads();
          for (int e = tid; e < WC * WC; e += JT) Gbuf[0][e % WC][e / WC] *= sc;
        }
        __syncthreads();
        if (tid == 0) { const uint64_t t1 = wall_clock64(); t_gram += t1 - t0; t0 = t1; }
        // ---- inner parallel Jacobi on G.  Every wave computes the parameters of all
        // 32 rotations of a round (lane l <-> pair l & 31) from the same LDS values, so
        // the waves agree without shared flags and a round needs a single barrier.
        bool visit_rot = false;
        const double tol2 = tol * tol;
        int cur = 0;
        for (int inner = 0; inner < MAX_INNER; ++inner) {
          bool any = false;
          for (int ir = 0; ir < WC - 1; ++ir) {
            const int pr = lane & (NPAIR - 1);
            int a = rr_player(pr, ir, WC), b = rr_player(WC - 1 - pr, ir, WC);
            if (a > b) { const int t = a; a = b; b = t; }
            double c = 1.0, s = 0.0;
            if (cols[a] >= 0 && cols[b] >= 0) {
              const double aa = Gbuf[cur][a][a], bb2 = Gbuf[cur][b][b], ab = Gbuf[cur][a][b];
              if (aa > 0.0 && bb2 > 0.0 && ab * ab > tol2 * aa * bb2) {
                const double dd = bb2 - aa;
                const double t = 2.0 * ab * copysign(1.0, dd) / (fabs(dd) + sqrt(dd * dd + 4.0 * ab * ab));
                c = rsqrt(1.0 + t * t);
                s = c * t;
              }
            }
            const bool round_any = __any(s != 0.0);
            any = any || round_any;
            if (round_any) {
              // G' = R^T G R, R = block diagonal of [[c, s], [-s, c]] on (a_k, b_k); each
              // 2x2 block (pair kr rows, pair kc columns) is owned by one thread, which
              // reads the current copy of G and writes the other one
              for (int e = tid; e < NPAIR * NPAIR; e += JT) {
                const int kr = e % NPAIR, kc = e / NPAIR;
                const double sr = __shfl(s, kr), sc = __shfl(s, kc);
                const double cr = __shfl(c, kr), cc = __shfl(c, kc);
                const int ar = __shfl(a, kr), br = __shfl(b, kr), ac = __shfl(a, kc), bc = __shfl(b, kc);
                const double g00 = Gbuf[cur][ar][ac], g01 = Gbuf[cur][ar][bc];
                const double g10 = Gbuf[cur][br][ac], g11 = Gbuf[cur][br][bc];
                const double t00 = cr * g00 - sr * g10, t01 = cr * g01 - sr * g11;
                const double t10 = sr * g00 + cr * g10, t11 = sr * g01 + cr * g11;
                Gbuf[cur ^ 1][ar][ac] = t00 * cc - t01 * sc;
                Gbuf[cur ^ 1][ar][bc] = t00 * sc + t01 * cc;
                Gbuf[cur ^ 1][br][ac] = t10 * cc - t11 * sc;
                Gbuf[cur ^ 1][br][bc] = t10 * sc + t11 * cc;
              }
              // V <- V R: row x, pair l
              for (int e = tid; e < WC * NPAIR; e += JT) {
                const int x = e % WC, l = e / WC;
                const double sl = __shfl(s, l), cl = __shfl(c, l);
                const int al = __shfl(a, l), bl = __shfl(b, l);
                if (sl == 0.0) continue;
                const double va = V[x][al], vb = V[x][bl];
                V[x][al] = cl * va - sl * vb;
                V[x][bl] = sl * va + cl * vb;
              }
              __syncthreads();
              cur ^= 1;
            }
          }
          visit_rot = visit_rot || any;
          if (!any) break;
        }
        if (tid == 0) { const uint64_t t1 = wall_clock64(); t_inner += t1 - t0; t0 = t1; }
        if (!visit_rot) continue;
        sweep_rot = true;
        // ---- apply V to the visited columns of B and J: C <- C V (wave w: row block w>>2, column block w&3)
        for (int which = 0; which < 2; ++which) {
          double* M = which == 0 ? S.B : S.J;
          const int64_t ld = which == 0 ? S.ldb : S.ldj;
          constexpr int PER = RCH * WC / JT;
          double pre[PER];
#pragma unroll
          for (int q = 0; q < PER; ++q) {
            const int e = tid + q * JT, rr = e % RCH, gc = cols[e / RCH];
            pre[q] = (gc >= 0 && rr < p) ? M[rr + int64_t(gc) * ld] : 0.0;
          }
          for (int r0 = 0; r0 < p; r0 += RCH) {
#pragma unroll
            for (int q = 0; q < PER; ++q) {
              const int e = tid + q * JT;
              chunk[e % RCH][e / RCH] = pre[q];
            }
            __syncthreads();
            if (r0 + RCH < p) {
#pragma unroll
              for (int q = 0; q < PER; ++q) {
                const int e = tid + q * JT, rr = r0 + RCH + e % RCH, gc = cols[e / RCH];
                pre[q] = (gc >= 0 && rr < p) ? M[rr + int64_t(gc) * ld] : 0.0;
              }
            }
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
        if (tid == 0) t_apply += wall_clock64() - t0;
      }
    }
    if (!sweep_rot) break;
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
  if (tid == 0 && S.info) {
    S.info[0] = sweeps_done;
    // phase times in units of 1024 ticks of the constant-rate wall clock
    S.info[1] = int32_t(t_gram >> 10);
    S.info[2] = int32_t(t_inner >> 10);
    S.info[3] = int32_t(t_apply >> 10);
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
    fl += 9.0 * double(j.p) * j.p * j.p;  // one sweep of plain one-sided Jacobi (nominal)
    by += 16.0 * double(j.p) * j.p;
  }
  ProfScope ps(ctx, "jacobi_svd", fl, by);
  hipLaunchKernelGGL(jacobi_block_kernel, dim3(unsigned(live.size())), dim3(JT), 0, ctx.stream, d);
  TTR_HIP(hipGetLastError());
}

}  // namespace ttr
