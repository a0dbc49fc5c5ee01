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
  __shared__ double pdiag[2][WC];   // diagonal of G for the next round's parameters
  __shared__ double pab[2][NPAIR];  // G(a_l, b_l) of the next round's pairs
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
  if (tid < WC) pdiag[0][tid] = G[tid][tid];
  if (tid < NPAIR) {
    const int a = rr_player(tid, 0, WC), b = rr_player(WC - 1 - tid, 0, WC);
    pab[0][tid] = G[min(a, b)][max(a, b)];
  }
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
  if (!__syncthreads_or(need)) {
    if (tid == 0 && S.info) {
      atomicAdd(reinterpret_cast<unsigned*>(S.info) + 1, unsigned(t1 - t0));
      atomicAdd(reinterpret_cast<unsigned*>(S.info) + 4, 1u);
    }
    return;
  }

  // ---- one inner Jacobi sweep on G (63 rounds of 32 disjoint rotations,
  // round-robin order).  A round has two phases:
  //   1. wave 0 computes the 32 rotation parameters from pdiag/pab and stores
  //      them in pc/ps, while every thread applies the previous round's
  //      rotation to its part of V (row lane, column slots 2w, 2w+1, 2w+16,
  //      2w+17 of wave w);
  //   2. thread (wave w, lane l) rotates its 2x2 blocks of G in place (row slot
  //      l & 31, column slots 2w + h and 2w + 16 + h, h = l >> 5) and stores the
  //      entries of G that the next round's parameters need into pdiag/pab.
  // Each phase ends with a barrier.  Every slot keeps its position in the
  // round-robin table, so the player at a position advances by one per round
  // (63 wraps to 1, player 0 stays at position 0); the positions a player
  // takes in the next round, and the partner there, are fixed per lane.
  const int h = lane >> 5;
  const int sl = lane & (NPAIR - 1);
  __shared__ double pc[NPAIR], ps[NPAIR];
  auto adv = [](int x) { return x == 0 ? 0 : (x == IR ? 1 : x + 1); };
  // row slot sl: players at positions sl and 63 - sl; in the next round the
  // player at position sl moves to nA (partner position 63 - nA, slot snA), the
  // one at 63 - sl to 62 - sl (partner position sl + 1, slot snB)
  int xA = rr_player(sl, 0, WC), xB = rr_player(IR - sl, 0, WC);
  const int nA = sl == 0 ? 0 : (sl == 1 ? IR : sl - 1);
  const int snA = min(nA, IR - nA), snB = min(62 - sl, sl + 1);
  int yA = rr_player(IR - nA, 1, WC), yB = rr_player(sl + 1, 1, WC);
  int xc[2], yc[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int kc = 2 * wave + 16 * j + h;
    xc[j] = rr_player(kc, 0, WC);
    yc[j] = rr_player(IR - kc, 0, WC);
  }
  const int qs[4] = {2 * wave, 2 * wave + 1, 2 * wave + 16, 2 * wave + 17};
  int xv[4], yv[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    xv[i] = rr_player(qs[i], 0, WC);
    yv[i] = rr_player(IR - qs[i], 0, WC);
  }
  double cV[4] = {1.0, 1.0, 1.0, 1.0}, sV[4] = {0.0, 0.0, 0.0, 0.0};
  int aV[4] = {0, 0, 0, 0}, bV[4] = {0, 0, 0, 0};
  bool visit_rot = false;
  for (int ir = 0; ir < IR; ++ir) {
    const int buf = ir & 1;
    const int a = min(xA, xB), b = max(xA, xB);
    // ---- phase 1
    if (wave == 0 && lane < NPAIR) {
      const double aa = pdiag[buf][a], bb2 = pdiag[buf][b], ab = pab[buf][sl];
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
      }
      pc[sl] = c;
      ps[sl] = s;
    }
    if (ir > 0) {
      double va[4], vb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        va[i] = V[lane][aV[i]];
        vb[i] = V[lane][bV[i]];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (sV[i] == 0.0) continue;
        V[lane][aV[i]] = cV[i] * va[i] - sV[i] * vb[i];
        V[lane][bV[i]] = sV[i] * va[i] + cV[i] * vb[i];
      }
    }
    __syncthreads();
    // ---- phase 2: all loads first, the rotations applied unconditionally (an
    // identity rotation reproduces its operands exactly), stores only where a
    // rotation acts
    const double c = pc[sl], s = ps[sl];
    visit_rot = visit_rot || __any(s != 0.0);
    const bool nxt = ir + 1 < IR;
    double cc[2], scc[2], g[2][4];
    int ac[2], bc[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int kc = 2 * wave + 16 * j + h;
      cc[j] = pc[kc];
      scc[j] = ps[kc];
      ac[j] = min(xc[j], yc[j]);
      bc[j] = max(xc[j], yc[j]);
      g[j][0] = G[a][ac[j]]; g[j][1] = G[a][bc[j]];
      g[j][2] = G[b][ac[j]]; g[j][3] = G[b][bc[j]];
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const double t00 = c * g[j][0] - s * g[j][2], t01 = c * g[j][1] - s * g[j][3];
      const double t10 = s * g[j][0] + c * g[j][2], t11 = s * g[j][1] + c * g[j][3];
      g[j][0] = t00 * cc[j] - t01 * scc[j];
      g[j][1] = t00 * scc[j] + t01 * cc[j];
      g[j][2] = t10 * cc[j] - t11 * scc[j];
      g[j][3] = t10 * scc[j] + t11 * cc[j];
      if (s != 0.0 || scc[j] != 0.0) {
        G[a][ac[j]] = g[j][0]; G[a][bc[j]] = g[j][1];
        G[b][ac[j]] = g[j][2]; G[b][bc[j]] = g[j][3];
      }
      if (nxt) {
        const int kc = 2 * wave + 16 * j + h;
        if (kc == sl) {  // diagonal block: G(a,a), G(b,b)
          pdiag[buf ^ 1][a] = g[j][0];
          pdiag[buf ^ 1][b] = g[j][3];
        }
        // next-round pairs (xA, yA) and (xB, yB): the thread holding row min and
        // column max of a pair stores its entry
        if (xA < yA && (yA == ac[j] || yA == bc[j]))
          pab[buf ^ 1][snA] = xA == a ? (yA == ac[j] ? g[j][0] : g[j][1]) : (yA == ac[j] ? g[j][2] : g[j][3]);
        if (xB < yB && (yB == ac[j] || yB == bc[j]))
          pab[buf ^ 1][snB] = xB == a ? (yB == ac[j] ? g[j][0] : g[j][1]) : (yB == ac[j] ? g[j][2] : g[j][3]);
      }
    }
    // this round's V rotation is applied in phase 1 of the next round
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      cV[i] = pc[qs[i]];
      sV[i] = ps[qs[i]];
      aV[i] = min(xv[i], yv[i]);
      bV[i] = max(xv[i], yv[i]);
      xv[i] = adv(xv[i]);
      yv[i] = adv(yv[i]);
    }
    xA = adv(xA); xB = adv(xB); yA = adv(yA); yB = adv(yB);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      xc[j] = adv(xc[j]);
      yc[j] = adv(yc[j]);
    }
    __syncthreads();
  }
  {
    double va[4], vb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      va[i] = V[lane][aV[i]];
      vb[i] = V[lane][bV[i]];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (sV[i] == 0.0) continue;
      V[lane][aV[i]] = cV[i] * va[i] - sV[i] * vb[i];
      V[lane][bV[i]] = sV[i] * va[i] + cV[i] * vb[i];
    }
  }
  __syncthreads();
  if (tid == 0) {
    t2 = wall_clock64();
    if (S.info) {
      atomicAdd(reinterpret_cast<unsigned*>(S.info) + 1, unsigned(t1 - t0));
      atomicAdd(reinterpret_cast<unsigned*>(S.info) + 2, unsigned(t2 - t1));
      atomicAdd(reinterpret_cast<unsigned*>(S.info) + 4, 1u);
    }
  }
  if (!visit_rot) return;
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
