// Batched one-sided (Hestenes) Jacobi SVD in FP64, block variant.
//
// Columns of each p x p matrix B are grouped in blocks of 32.  A sweep visits
// every pair of blocks once in round-robin order: each round is one launch in
// which every disjoint block pair of every matrix is one workgroup, so a large
// matrix is worked on by many CUs at once and a batch of uneven matrices keeps
// the whole GPU busy.  A visit
//   1. forms the Gram matrix G = C^T C of the (up to) 64 columns C of the
//      block pair in LDS with FP64 MFMA, straight from the current columns
//      (rotation angles always come from fresh inner products, as in plain
//      one-sided Jacobi);
//   2. runs one parallel cyclic Jacobi sweep on G in LDS (32 disjoint rotations
//      per round, round-robin order; each 2x2 block of G and each row pair of V
//      is updated in place by one thread), accumulating the rotation V; a pair
//      is rotated when |G_ij| > tol sqrt(G_ii G_jj), tol = p eps;
//   3. if anything rotated, applies V to the visited columns of B and of the
//      accumulated right rotations J with FP64 MFMA (C <- C V, 32-row chunks
//      staged in LDS) and marks the matrix as not yet converged.
// A matrix whose previous sweep rotated nothing is converged: its workgroups
// return at once.  On exit B_out = B_in J has orthogonal columns,
// B_out[:,c] = u_c s_c, and sv[c] = |B_out[:,c]|.
#include "ttr_internal.h"

namespace ttr {

namespace {

typedef double d4v __attribute__((ext_vector_type(4)));

constexpr int JT = 512;         // threads per visit (8 waves)
constexpr int WB = 32;          // column block width
constexpr int WC = 2 * WB;      // columns per visit
constexpr int GLD = WC + 1;     // padded LDS row
constexpr int RCH = 32;         // row chunk for Gram / apply
constexpr int NPAIR = WC / 2;
constexpr int IR = WC - 1;      // rounds of one inner sweep
constexpr int MAX_SWEEPS = 40;

struct JacJob {
  double* B;
  int64_t ldb;
  double* J;
  int64_t ldj;
  int32_t* info;
  int p, nblk, pp, rounds, slot0;
};

__host__ __device__ __forceinline__ int rr_player(int pos, int r, int pp) {
  return pos == 0 ? 0 : ((pos - 1 + r) % (pp - 1)) + 1;
}
// 1/sqrt(x) and 1/x for positive normal x: the hardware estimate refined by two
// Newton-Raphson steps, y += y (1 - x y^2) / 2 and y += y (1 - x y)
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
  }
  return y;
}

// lane 0's value broadcast through scalar registers (the readfirstlane builtin
// takes a 32-bit int: a double passed to it directly is converted, not copied)
__device__ __forceinline__ double first_lane(double v) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(b & 0xffffffffu)));
  const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(b >> 32)));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// ---- one inner Jacobi sweep on G: 63 rounds of 32 disjoint rotations in
// round-robin order (pair t of round r joins the players at positions t and
// 63 - t; every player but 0 advances one position per round).  G is
// symmetric and only its upper triangle is read or written from here on:
// entry (i, j) lives at G[min(i,j)][max(i,j)].  A round has two phases, each
// ending with a barrier:
//   1. lanes 0..31 of wave 0 compute the rotation (c, s) of pair `lane` from
//      G into pcs[r & 1]; waves 1..7 meanwhile apply the previous round's
//      rotations to V (row = lane; wave w owns pairs w - 1 + 7k), skipping
//      identity rotations (their parameters are wave-uniform, read by one
//      lane and broadcast through scalar registers);
//   2. every thread rotates its upper-triangle 2x2 block(s) of G in place:
//      threads 0..495 one off-diagonal block (pair positions P < Q), threads
//      496..511 two diagonal blocks (P = Q).  A block whose two rotations
//      are the identity is neither read nor written.
// Each thread only touches its own blocks in phase 2, and every block is
// owned by one thread, so the update is in place.
__device__ __forceinline__ void inner_sweep(double (*G)[GLD], double (*V)[GLD], double tol2, int tid, int* vrot) {
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  __shared__ double2 pcs[2][NPAIR];
  auto adv = [](int x) { return x == 0 ? 0 : (x == IR ? 1 : x + 1); };
  // phase-1 players of the parameter lanes
  int qa = rr_player(lane & (NPAIR - 1), 0, WC), qb = rr_player(IR - (lane & (NPAIR - 1)), 0, WC);
  // phase-2 blocks
  int bP[2], bQ[2];
  const int nbk = tid < 496 ? 1 : 2;
  if (tid < 496) {
    int t = tid, P = 0;
    while (t >= NPAIR - 1 - P) {
      t -= NPAIR - 1 - P;
      ++P;
    }
    bP[0] = P;
    bQ[0] = P + 1 + t;
    bP[1] = bQ[1] = 0;
  } else {
    bP[0] = bQ[0] = 2 * (tid - 496);
    bP[1] = bQ[1] = 2 * (tid - 496) + 1;
  }
  int rx0[2], rx1[2], cx0[2], cx1[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    rx0[b] = rr_player(bP[b], 0, WC);
    rx1[b] = rr_player(IR - bP[b], 0, WC);
    cx0[b] = rr_player(bQ[b], 0, WC);
    cx1[b] = rr_player(IR - bQ[b], 0, WC);
  }
  // V pairs of this wave (waves 1..7), players one round behind
  constexpr int VK = 5;
  int vt[VK], vx[VK], vy[VK];
#pragma unroll
  for (int k = 0; k < VK; ++k) {
    vt[k] = wave - 1 + 7 * k;
    const int t = (wave > 0 && vt[k] < NPAIR) ? vt[k] : 0;
    vx[k] = rr_player(t, 0, WC);
    vy[k] = rr_player(IR - t, 0, WC);
  }
  // (c, s) of this wave's V pairs for the pending V update, read at the end of
  // the phase 2 that follows their computation
  double vc[VK], vs[VK];
#pragma unroll
  for (int k = 0; k < VK; ++k) {
    vc[k] = 1.0;
    vs[k] = 0.0;
  }
  auto v_params = [&](const double2* cs) {
#pragma unroll
    for (int k = 0; k < VK; ++k) {
      if (vt[k] < NPAIR) {
        const double2 cv = cs[vt[k]];
        vc[k] = cv.x;
        vs[k] = cv.y;
      }
    }
  };
  // all loads first, then the rotations, stores only where a rotation acts
  auto v_update = [&]() {
    double va[VK], vb[VK];
    int lo[VK], hi[VK];
#pragma unroll
    for (int k = 0; k < VK; ++k) {
      lo[k] = min(vx[k], vy[k]);
      hi[k] = max(vx[k], vy[k]);
      va[k] = V[lane][lo[k]];
      vb[k] = V[lane][hi[k]];
    }
#pragma unroll
    for (int k = 0; k < VK; ++k) {
      if (vt[k] < NPAIR && vs[k] != 0.0) {
        V[lane][lo[k]] = vc[k] * va[k] - vs[k] * vb[k];
        V[lane][hi[k]] = vs[k] * va[k] + vc[k] * vb[k];
      }
      vx[k] = adv(vx[k]);
      vy[k] = adv(vy[k]);
    }
  };
  for (int ir = 0; ir < IR; ++ir) {
    const int buf = ir & 1;
    // ---- phase 1
    if (wave == 0) {
      if (lane < NPAIR) {
        const int a = min(qa, qb), b = max(qa, qb);
        const double aa = G[a][a], bb2 = G[b][b], ab = G[a][b];
        double c = 1.0, s = 0.0;
        // padding columns have a zero diagonal and never rotate
        if (aa > 0.0 && bb2 > 0.0 && ab * ab > tol2 * aa * bb2) {
          const double dd = bb2 - aa;
          const double q = dd * dd + 4.0 * ab * ab;
          double t;
          if (q > 0x1p-900) {
            // G is scaled to max diagonal ~1, so q <= ~20: hardware estimates with two
            // Newton steps each (rsqrt, reciprocal) are accurate to ~1 ulp here
            const double rq = rsqrt_nr(q);
            t = 2.0 * ab * copysign(1.0, dd) * rcp_nr(fabs(dd) + q * rq);
          } else {
            t = 2.0 * ab * copysign(1.0, dd) / (fabs(dd) + sqrt(q));
          }
          c = rsqrt_nr(1.0 + t * t);  // 1 + t^2 in [1, 2]
          s = c * t;
          *vrot = 1;
        }
        pcs[buf][lane] = make_double2(c, s);
        qa = adv(qa);
        qb = adv(qb);
      }
    } else if (ir > 0) {
      v_update();
    }
    __syncthreads();
    // ---- phase 2: loads first (parameters and the four entries), then the
    // rotation, stores only where a rotation acts
    for (int b = 0; b < nbk; ++b) {
      const int rl = min(rx0[b], rx1[b]), rh = max(rx0[b], rx1[b]);
      const int cl = min(cx0[b], cx1[b]), ch = max(cx0[b], cx1[b]);
      double* e00 = &G[min(rl, cl)][max(rl, cl)];
      double* e01 = &G[min(rl, ch)][max(rl, ch)];
      double* e10 = &G[min(rh, cl)][max(rh, cl)];
      double* e11 = &G[min(rh, ch)][max(rh, ch)];
      const double2 pr = pcs[buf][bP[b]], pq = pcs[buf][bQ[b]];
      const double g0 = *e00, g1 = *e01, g2 = *e10, g3 = *e11;
      const double t00 = pr.x * g0 - pr.y * g2, t01 = pr.x * g1 - pr.y * g3;
      const double t10 = pr.y * g0 + pr.x * g2, t11 = pr.y * g1 + pr.x * g3;
      if (pr.y != 0.0 || pq.y != 0.0) {
        *e00 = t00 * pq.x - t01 * pq.y;
        *e01 = t00 * pq.y + t01 * pq.x;
        if (bP[b] != bQ[b]) *e10 = t10 * pq.x - t11 * pq.y;  // a diagonal block's (h, l) is its (l, h)
        *e11 = t10 * pq.y + t11 * pq.x;
      }
      rx0[b] = adv(rx0[b]);
      rx1[b] = adv(rx1[b]);
      cx0[b] = adv(cx0[b]);
      cx1[b] = adv(cx1[b]);
    }
    if (wave > 0) v_params(pcs[buf]);
    __syncthreads();
  }
  if (wave > 0) v_update();
  __syncthreads();
}

__global__ __launch_bounds__(JT, 2) void jacobi_round_kernel(const JacJob* __restrict__ jobs, int njobs, int round,
                                                             int sweep, int32_t* __restrict__ flags) {
  const int bid = blockIdx.x;
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].slot0 <= bid) lo = mid; else hi = mid - 1;
  }
  const JacJob S = jobs[lo];
  if (round >= S.rounds) return;
  if (sweep > 0 && flags[((sweep - 1) & 1) * njobs + lo] == 0) return;  // converged
  const int k = bid - S.slot0;
  const int p = S.p;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (k == 0 && round == 0 && tid == 0 && S.info) S.info[0] = sweep + 1;
  uint64_t t0 = 0, t1 = 0, t2 = 0;
  if (tid == 0) t0 = wall_clock64();

  int ba, bb;
  if (S.nblk == 1) { ba = 0; bb = -1; }
  else {
    ba = rr_player(k, round, S.pp);
    bb = rr_player(S.pp - 1 - k, round, S.pp);
    if (ba >= S.nblk) ba = -1;
    if (bb >= S.nblk) bb = -1;
    if (ba < 0 && bb < 0) return;
  }

  // G doubles as the row-chunk staging buffer of the Gram and apply phases
  __shared__ double G[WC][GLD];
  __shared__ double V[WC][GLD];
  __shared__ int cols[WC];          // global column of local index, -1 if padding
  double (*chunk)[GLD] = G;

  if (tid < WC) {
    const int blk = tid < WB ? ba : bb;
    const int c = blk < 0 ? -1 : blk * WB + (tid % WB);
    cols[tid] = (c >= 0 && c < p) ? c : -1;
  }
  __syncthreads();

  // ---- Gram matrix G = C^T C (MFMA f64 16x16x4; wave w owns output blocks w and w+8)
  constexpr int PER = RCH * WC / JT;  // chunk elements per thread
  {
    d4v gacc[2] = {d4v{0, 0, 0, 0}, d4v{0, 0, 0, 0}};
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
      for (int q = 0; q < 4; ++q) G[16 * bi + 4 * q + (lane >> 4)][16 * bj + (lane & 15)] = gacc[t][q];
    }
  }
  for (int e = tid; e < WC * WC; e += JT) V[e % WC][e / WC] = (e % WC == e / WC) ? 1.0 : 0.0;
  __syncthreads();
  if (tid == 0) t1 = wall_clock64();
  // scale G by a power of two (rotation angles are scale free) so that the
  // squares below cannot overflow for unnormalised inputs
  {
    double mx = 0.0;
    for (int i = lane; i < WC; i += 64) mx = fmax(mx, G[i][i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
    const double sc = (mx > 0.0 && isfinite(mx)) ? ldexp(1.0, -ilogb(mx)) : 1.0;
    __syncthreads();
    for (int e = tid; e < WC * WC; e += JT) G[e % WC][e / WC] *= sc;
  }
  __syncthreads();
  const double tol = double(p > 16 ? p : 16) * 2.220446049250313e-16;
  const double tol2 = tol * tol;
  // a visit whose columns are already pairwise orthogonal changes nothing (no
  // round would rotate): skip the inner sweep and the apply
  int need = 0;
  for (int e = tid; e < WC * WC; e += JT) {
    const int i = e % WC, j = e / WC;
    if (i < j) {
      const double aa = G[i][i], bb2 = G[j][j], ab = G[i][j];
      need |= (aa > 0.0 && bb2 > 0.0 && ab * ab > tol2 * aa * bb2) ? 1 : 0;
    }
  }
  __shared__ int vrot;
  if (tid == 0) vrot = 0;
  if (!__syncthreads_or(need)) {
    if (tid == 0 && S.info) {
      atomicAdd(reinterpret_cast<unsigned*>(S.info) + 1, unsigned(t1 - t0));
      atomicAdd(reinterpret_cast<unsigned*>(S.info) + 4, 1u);
    }
    return;
  }

  inner_sweep(G, V, tol2, tid, &vrot);
  if (tid == 0) {
    t2 = wall_clock64();
    if (S.info) {
      atomicAdd(reinterpret_cast<unsigned*>(S.info) + 1, unsigned(t1 - t0));
      atomicAdd(reinterpret_cast<unsigned*>(S.info) + 2, unsigned(t2 - t1));
      atomicAdd(reinterpret_cast<unsigned*>(S.info) + 4, 1u);
    }
  }
  if (!vrot) return;
  if (tid == 0) {
    flags[(sweep & 1) * njobs + lo] = 1;
    if (S.info) atomicAdd(reinterpret_cast<unsigned*>(S.info) + 5, 1u);
  }

  // ---- apply V to the visited columns of B and J: C <- C V (wave w: row block w>>2, column block w&3)
  for (int which = 0; which < 2; ++which) {
    double* M = which == 0 ? S.B : S.J;
    const int64_t ld = which == 0 ? S.ldb : S.ldj;
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
        const double av = chunk[16 * bi + (lane & 15)][kk + (lane >> 4)];
        const double bv = V[kk + (lane >> 4)][16 * bj + (lane & 15)];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
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
  if (tid == 0 && S.info) atomicAdd(reinterpret_cast<unsigned*>(S.info) + 3, unsigned(wall_clock64() - t2));
}

// singular values: column norms of B (one workgroup per matrix, one wave per column)
__global__ __launch_bounds__(256) void column_norms_kernel(const SVDJob* __restrict__ jobs) {
  const SVDJob S = jobs[blockIdx.x];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c = wave; c < S.p; c += 4) {
    const double* x = S.B + int64_t(c) * S.ldb;
    double s = 0.0;
    for (int i = lane; i < S.p; i += 64) s += x[i] * x[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) S.sv[c] = sqrt(s);
  }
}

}  // namespace

void jacobi_svd_batched(Ctx& ctx, std::vector<SVDJob>& jobs) {
  std::vector<SVDJob> live;
  for (auto& j : jobs)
    if (j.p > 0) live.push_back(j);
  if (live.empty()) return;
  const int nj = int(live.size());
  std::vector<JacJob> jj(nj);
  std::vector<Fill2D> fills;
  int slots = 0, max_rounds = 1;
  for (int q = 0; q < nj; ++q) {
    const SVDJob& s = live[q];
    JacJob& J = jj[q];
    J.B = s.B; J.ldb = s.ldb; J.J = s.J; J.ldj = s.ldj; J.info = s.info; J.p = s.p;
    J.nblk = (s.p + WB - 1) / WB;
    J.pp = J.nblk + (J.nblk & 1);
    J.rounds = J.nblk == 1 ? 1 : J.pp - 1;
    J.slot0 = slots;
    slots += J.nblk == 1 ? 1 : J.pp / 2;
    max_rounds = std::max(max_rounds, J.rounds);
    fills.push_back(Fill2D{s.J, s.ldj, s.p, s.p, 1.0});
  }
  const JacJob* dj = ctx.upload(jj);
  const SVDJob* dl = ctx.upload(live);
  int32_t* flags = ctx.ws.alloc_n<int32_t>(size_t(2) * nj);
  int32_t* hflags = static_cast<int32_t*>(ctx.pin.alloc(sizeof(int32_t) * nj));
  fill2d_batched(ctx, fills);
  zero_fill(ctx, flags, sizeof(int32_t) * 2 * nj);
  for (int sweep = 0; sweep < MAX_SWEEPS; ++sweep) {
    int32_t* cur = flags + size_t(sweep & 1) * nj;
    if (sweep >= 2) zero_fill(ctx, cur, sizeof(int32_t) * nj);
    for (int r = 0; r < max_rounds; ++r) {
      // timed per launch; the work depends on the sweeps taken, so the caller
      // adds it from the counters
      ProfScope ps(ctx, "jacobi_svd");
      hipLaunchKernelGGL(jacobi_round_kernel, dim3(unsigned(slots)), dim3(JT), 0, ctx.stream, dj, nj, r, sweep, flags);
      TTR_HIP(hipGetLastError());
    }
    // convergence check from the third sweep on (no matrix converges earlier in practice)
    if (sweep >= 2) {
      TTR_HIP(hipMemcpyAsync(hflags, cur, sizeof(int32_t) * nj, hipMemcpyDeviceToHost, ctx.stream));
      ctx.sync();
      bool any = false;
      for (int q = 0; q < nj && !any; ++q) any = hflags[q] != 0;
      if (!any) break;
    }
  }
  hipLaunchKernelGGL(column_norms_kernel, dim3(unsigned(nj)), dim3(256), 0, ctx.stream, dl);
  TTR_HIP(hipGetLastError());
}

}  // namespace ttr
