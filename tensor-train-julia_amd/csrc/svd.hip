// Batched one-sided (Hestenes) Jacobi SVD in FP64, block variant.
//
// Columns of each p x p matrix B are grouped in blocks of 32 (or 16).  A sweep visits
// every pair of blocks once in round-robin order: each round is one launch in
// which every disjoint block pair of every matrix is one workgroup, so a large
// matrix is worked on by many CUs at once and a batch of uneven matrices keeps
// the whole GPU busy.  A visit
//   1. forms the upper triangle of the Gram matrix G = C^T C of the (up to) 64
//      columns C of the block pair in LDS with FP64 MFMA, straight from the
//      current columns (rotation angles always come from fresh inner products,
//      as in plain one-sided Jacobi);
//   2. runs one parallel cyclic Jacobi sweep on G (32 disjoint rotations per
//      round, round-robin order, G in LDS and the accumulated rotation V in
//      registers; see inner_sweep); a pair is rotated when
//      |G_ij| > tol sqrt(G_ii G_jj), tol = p eps (a relative test: the small
//      singular values stay relatively accurate, which the pseudo-inverses of
//      STTA need);
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

constexpr int RCH = 32;         // row chunk for Gram / apply
constexpr int RING = 3;         // row chunks in flight (registers) ahead of the MFMAs
constexpr int MAX_SWEEPS = 40;
constexpr int JAC_MAX_SPLIT = 8;  // row ranges per visit in the split launches

// Shape of a visit for column blocks of width WB_ (16 or 32)
template <int WB_>
struct Jac {
  static constexpr int WB = WB_;             // column block width
  static constexpr int WC = 2 * WB;          // columns per visit
  static constexpr int GLD = WC + 2;         // LDS row pitch (2 mod 32 doubles: see inner_sweep)
  static constexpr int JT = 16 * WB;         // threads per visit
  static constexpr int NW = JT / 64;         // waves
  static constexpr int NPAIR = WB;           // rotations per inner round
  static constexpr int IR = WC - 1;          // rounds of one inner sweep
  static constexpr int NT = WC / 16;         // 16x16 MFMA tiles per Gram row
  static constexpr int TPW = NT * NT / NW;   // Gram tiles per wave (at most)
  // The inner sweep reads only the upper triangle of G: the NGT tiles bi <= bj.
  // Tile k goes to wave k % NW; its (bi, bj) are bits 4k.. and 4k+2.. of GTILES.
  // For 4 x 4 tiles the order puts (0,3),(1,3) on wave 0 and (2,3),(3,3) on
  // wave 1, so a wave with two tiles reads one column operand for both.
  static constexpr int NGT = NT * (NT + 1) / 2;
  static constexpr unsigned long long GTILES =
      NT == 4 ? (0x0ull | (0x3ull << 2)) | ((0x2ull | (0x3ull << 2)) << 4) | ((0x0ull | (0x0ull << 2)) << 8) |
                    ((0x0ull | (0x1ull << 2)) << 12) | ((0x1ull | (0x1ull << 2)) << 16) |
                    ((0x0ull | (0x2ull << 2)) << 20) | ((0x1ull | (0x2ull << 2)) << 24) |
                    ((0x2ull | (0x2ull << 2)) << 28) | ((0x1ull | (0x3ull << 2)) << 32) |
                    ((0x3ull | (0x3ull << 2)) << 36)
              : (0x0ull | (0x0ull << 2)) | ((0x0ull | (0x1ull << 2)) << 4) | ((0x1ull | (0x1ull << 2)) << 8);
  static constexpr int VROWS = WB / 8;       // rows of V per thread
  static constexpr int PER = RCH * WC / JT;  // staged chunk elements per thread
  static constexpr int MINB = WB == 16 ? 4 : 2;  // resident visits per CU the registers allow
};

struct JacJob {
  double* B;
  int64_t ldb;
  double* J;
  int64_t ldj;
  int32_t* info;
  double tol;  // rotate a pair when |G_ij| > tol sqrt(G_ii G_jj)
  int p, nblk, pp, rounds, slot0;
  int nsplit;   // row ranges of a visit's Gram and apply in the split launches
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

// Cross-lane moves of a double on the VALU (DPP): lane i takes lane i - 1's
// value (wave_shr:1; lane 0 gets 0) or lane i + 1's (wave_shl:1; lane 63 gets 0).
// bound_ctrl set: the lane without a source writes 0 by itself, so no
// destination register has to be cleared first.
__device__ __forceinline__ double from_lower_lane(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_upper_lane(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

// ---- one inner Jacobi sweep on the 64x64 Gram matrix: 63 rounds of 32
// disjoint rotations in round-robin order.  Pair t of a round joins positions
// t and 63 - t; after the round every value moves with its player, position p
// to adv(p) (0 stays, 63 wraps to 1, others p + 1), so a player's data sits at
// the position it takes in the next round, and after 63 rounds every player is
// back at its starting position.  Working in these positions keeps every
// thread's addresses fixed:
//   * G (symmetric, upper triangle only) ping-pongs between two LDS buffers:
//     every 2x2 block of G (rows of one pair, columns of another) is owned by
//     one thread, read at its positions in the current buffer and written
//     moved into the other, at addresses fixed for the whole sweep;
//   * V lives in registers: the same thread holds V at rows 2w + (l >> 5) + 16j
//     (j = 0..3) and positions P, 63 - P; the move is a shift by one lane
//     within each 32-lane half (DPP, plus the two ends of the round-robin
//     table), so V never touches LDS during the sweep.
// Lanes 0..31 of wave 0 compute pair t's rotation (c, s) (position t first);
// the phases of a round are described at the loop.  On exit V, in player order
// again, is stored in GA (GB is free); *vrot is set when any rotation acted.
template <int WB_>
__device__ __forceinline__ void inner_sweep(double (*GA)[Jac<WB_>::GLD], double (*GB)[Jac<WB_>::GLD], double tol2,
                                            int tid, int* vrot) {
  using C = Jac<WB_>;
  constexpr int GLD = C::GLD, NPAIR = C::NPAIR, IR = C::IR;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int P = lane & (NPAIR - 1), slot = tid / NPAIR;
  __shared__ double2 pcs[NPAIR];
  auto adv = [](int x) { return x == 0 ? 0 : (x == IR ? 1 : x + 1); };
  const int x0 = P, x1 = IR - P;
  // G blocks of this thread in the upper-triangle storage: block (bP, bQ) holds the
  // rows of pair bP and the columns of pair bQ, entry (i, j) at [min][max].  Group g
  // of NPAIR lanes takes the blocks {P, P + g + 1 mod NPAIR} (g < NPAIR/2 - 1); the
  // last group the blocks {P, P + NPAIR/2} (P < NPAIR/2) and two diagonal blocks per
  // lane (P >= NPAIR/2); every unordered pair of pairs once.  With a row pitch of 2
  // mod 32 doubles the 32 lanes of a group touch 32 different bank pairs (up to the
  // wrap of P + g + 1) for each of the four entries.
  constexpr int HALF = NPAIR / 2;
  const int g = tid / NPAIR;
  int nbk = 0, bP[2] = {0, 0}, bQ[2] = {0, 0};
  if (g < HALF - 1) {
    nbk = 1;
    bP[0] = P;
    bQ[0] = (P + g + 1) & (NPAIR - 1);
  } else if (g == HALF - 1) {
    if (P < HALF) {
      nbk = 1;
      bP[0] = P;
      bQ[0] = P + HALF;
    } else {
      nbk = 2;
      bP[0] = bQ[0] = 2 * (P - HALF);
      bP[1] = bQ[1] = 2 * (P - HALF) + 1;
    }
  }
  auto canon = [](int i, int j) { return i < j ? i * GLD + j : j * GLD + i; };
  int rd[2][4], wr[2][4];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int r0 = bP[b], r1 = IR - bP[b], c0 = bQ[b], c1 = IR - bQ[b];
    rd[b][0] = canon(r0, c0); rd[b][1] = canon(r0, c1);
    rd[b][2] = canon(r1, c0); rd[b][3] = canon(r1, c1);
    wr[b][0] = canon(adv(r0), adv(c0)); wr[b][1] = canon(adv(r0), adv(c1));
    wr[b][2] = canon(adv(r1), adv(c0)); wr[b][3] = canon(adv(r1), adv(c1));
  }
  // V rows of this thread, identity at the start
  double VL[C::VROWS], VR[C::VROWS];
#pragma unroll
  for (int j = 0; j < C::VROWS; ++j) {
    const int row = slot + 16 * j;
    VL[j] = row == x0 ? 1.0 : 0.0;
    VR[j] = row == x1 ? 1.0 : 0.0;
  }
  // sources of the move (groups of NPAIR lanes): position t takes t - 1 (t >= 2:
  // the left slot of the lane below), IR (t = 1: the right slot of the group's
  // first lane, which therefore shifts its right slot instead of its left) or
  // itself (t = 0); position IR - t takes IR - 1 - t, the right slot of the lane
  // above, or the own left slot (t = NPAIR - 1).  The lanes that receive a
  // value from across a group boundary are exactly those that do not use it.
  // rotation parameters of one round from the G buffer g (lanes 0..NPAIR-1 of wave 0)
  auto rotations = [&](const double (*g)[GLD]) {
    if (lane < NPAIR) {
      const double aa = g[x0][x0], bb2 = g[x1][x1], ab = g[x0][x1];
      double c = 1.0, s = 0.0;
      // padding columns have a zero diagonal and never rotate (bitwise & keeps
      // the three loads ahead of a single wait)
      if ((aa > 0.0) & (bb2 > 0.0) & (ab * ab > tol2 * aa * bb2)) {
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
      pcs[lane] = make_double2(c, s);
    }
  };
  // one round's rotation and move of this thread's V rows
  auto v_update = [&](double2 pr) {
#pragma unroll
    for (int j = 0; j < C::VROWS; ++j) {
      const double nl = pr.x * VL[j] - pr.y * VR[j];
      const double nr = pr.y * VL[j] + pr.x * VR[j];
      const double fromL = from_lower_lane(P == 0 ? nr : nl);
      const double fromR = from_upper_lane(nr);
      VL[j] = P == 0 ? nl : fromL;
      VR[j] = P == NPAIR - 1 ? nl : fromR;
    }
  };
  // A round is two phases, each ending with a barrier.  Phase A: every thread
  // rotates its G blocks (current -> next buffer, moved).  Phase B: wave 0
  // computes the next round's rotations from the new buffer while the other
  // waves rotate their V rows; wave 0 rotates its own V rows one round late,
  // in the next phase A (its rows are touched by no other wave), so no wave
  // does both the rotation parameters and V work between the same barriers.
  if (wave == 0) rotations(GA);
  __syncthreads();
  double2 held = make_double2(1.0, 0.0);
  // one round from buffer cur into buffer nxt; the loop below is unrolled by two
  // so that cur and nxt are fixed arrays at each call site and every LDS address
  // is a loop-invariant register plus an immediate offset
  auto round_step = [&](const double (*cur)[GLD], double (*nxt)[GLD], int ir) {
    // ---- phase A
    const double2 pr = pcs[P];
    const double* cb = &cur[0][0];
    double* nb = &nxt[0][0];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      if (b < nbk) {
        const double2 pa = pcs[bP[b]], pq = pcs[bQ[b]];
        const double g00 = cb[rd[b][0]], g01 = cb[rd[b][1]], g10 = cb[rd[b][2]], g11 = cb[rd[b][3]];
        const double t00 = pa.x * g00 - pa.y * g10, t01 = pa.x * g01 - pa.y * g11;
        const double t10 = pa.y * g00 + pa.x * g10, t11 = pa.y * g01 + pa.x * g11;
        nb[wr[b][0]] = t00 * pq.x - t01 * pq.y;
        nb[wr[b][1]] = t00 * pq.y + t01 * pq.x;
        if (bP[b] != bQ[b]) nb[wr[b][2]] = t10 * pq.x - t11 * pq.y;  // a diagonal block's (h, l) is its (l, h)
        nb[wr[b][3]] = t10 * pq.y + t11 * pq.x;
      }
    }
    if (wave == 0 && ir > 0) v_update(held);
    __syncthreads();
    // ---- phase B
    if (wave == 0) {
      if (ir + 1 < IR) rotations(nxt);
      held = pr;
    } else {
      v_update(pr);
    }
    __syncthreads();
  };
  for (int ir = 0; ir < IR; ir += 2) {
    round_step(GA, GB, ir);
    if (ir + 1 < IR) round_step(GB, GA, ir + 1);
  }
  if (wave == 0) v_update(held);
  // V back in player order (positions = players after 63 moves) into GA
#pragma unroll
  for (int j = 0; j < C::VROWS; ++j) {
    const int row = slot + 16 * j;
    GA[row][x0] = VL[j];
    GA[row][x1] = VR[j];
  }
  __syncthreads();
}

// MODE 0: a whole visit in one workgroup.  The split launches spread a visit over
// several workgroups (a matrix's visits are few when the batch holds few
// matrices): MODE 1 forms the Gram of one row range of the pair's columns into
// a partial buffer (blockIdx.y = the range), MODE 2 adds the partials in range
// order, runs the inner sweep and stores V and whether it rotated, MODE 3 applies
// V to one row range of B and J.  With one range MODE 1-3 compute what MODE 0 does,
// bit for bit.
template <int WB_, int MODE>
__global__ __launch_bounds__(Jac<WB_>::JT, Jac<WB_>::MINB) void jacobi_round_kernel(
    const JacJob* __restrict__ jobs, int njobs, int round, int sweep, int32_t* __restrict__ flags,
    double* __restrict__ gpart, int gstride, double* __restrict__ vbuf, int32_t* __restrict__ rotf) {
  using C = Jac<WB_>;
  constexpr int WB = C::WB, WC = C::WC, GLD = C::GLD, JT = C::JT, NW = C::NW, NT = C::NT;
  const int bid = blockIdx.x, split = blockIdx.y;
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].slot0 <= bid) lo = mid; else hi = mid - 1;
  }
  const JacJob S = jobs[lo];
  if (round >= S.rounds || split >= S.nsplit) return;
  if (sweep > 0 && flags[((sweep - 1) & 1) * njobs + lo] == 0) return;  // converged
  const int k = bid - S.slot0;
  const int p = S.p;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (MODE != 1 && MODE != 3 && k == 0 && round == 0 && tid == 0 && S.info) S.info[0] = sweep + 1;
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

  // GA holds the Gram matrix (and stages the Gram's row chunks); the inner sweep
  // ping-pongs G between GA and GB and leaves V in GA; GB then stages the
  // apply's row chunks
  __shared__ double GA[WC][GLD];
  __shared__ double GB[WC][GLD];
  __shared__ int cols[WC];          // global column of local index, -1 if padding
  double (*G)[GLD] = GA;
  double (*chunk)[GLD] = GA;

  if (tid < WC) {
    const int blk = tid < WB ? ba : bb;
    const int c = blk < 0 ? -1 : blk * WB + (tid % WB);
    cols[tid] = (c >= 0 && c < p) ? c : -1;
  }
  __syncthreads();

  // ---- Gram matrix G = C^T C, upper-triangle tiles (MFMA f64 16x16x4)
  // Row chunks of RCH rows pass through LDS; the global loads run RING chunks
  // ahead in registers, so each load has RING chunks of MFMA work to arrive.
  constexpr int PER = C::PER;
  const int nch = (p + RCH - 1) / RCH;
  auto load_chunk = [&](const double* M, int64_t ld, int r0, double (&dst)[PER]) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + q * JT, rr = r0 + e % RCH, gc = cols[e / RCH];
      dst[q] = (gc >= 0 && rr < p) ? M[rr + int64_t(gc) * ld] : 0.0;
    }
  };
  auto stage_chunk = [&](double (*buf)[GLD], const double (&src)[PER]) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + q * JT;
      buf[e % RCH][e / RCH] = src[q];
    }
  };
  if (MODE == 2) {
    // the partial Grams of the row ranges, added in range order (upper-triangle tiles)
    // every load issued before the first add (the sums are latency-bound)
    const double* gp = gpart + size_t(bid) * gstride * WC * WC;
    constexpr int EPT = WC * WC / JT;
    double v[EPT][JAC_MAX_SPLIT];
#pragma unroll
    for (int u = 0; u < EPT; ++u) {
      const int e = tid + JT * u, i = e / WC, j = e % WC;
#pragma unroll
      for (int q = 0; q < JAC_MAX_SPLIT; ++q)
        v[u][q] = (q < S.nsplit && i / 16 <= j / 16) ? gp[size_t(q) * WC * WC + e] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < EPT; ++u) {
      const int e = tid + JT * u, i = e / WC, j = e % WC;
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < JAC_MAX_SPLIT; ++q)
        if (q < S.nsplit) acc += v[u][q];
      if (i / 16 <= j / 16) G[i][j] = acc;
    }
  } else if (MODE != 3) {
    // the Gram's chunk is staged column-major ([column][row], pitch 2 mod 32
    // doubles): the MFMA operands are then 16 columns x 4 rows per read, which
    // fall in distinct banks; the staging stores are contiguous
    constexpr int CP = RCH + 2;
    static_assert(WC * CP <= WC * GLD, "the Gram chunk must fit in GA");
    // two chunk buffers (GA, GB; GB is free until the inner sweep): chunk c + 1 is
    // staged while the MFMAs of chunk c run, one barrier per chunk
    double (*gbuf0)[CP] = reinterpret_cast<double (*)[CP]>(&GA[0][0]);
    double (*gbuf1)[CP] = reinterpret_cast<double (*)[CP]>(&GB[0][0]);
    auto gstage = [&](double (*buf)[CP], const double (&src)[PER]) {
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int e = tid + q * JT;
        buf[e / RCH][e % RCH] = src[q];
      }
    };
    d4v gacc[C::TPW];
#pragma unroll
    for (int t = 0; t < C::TPW; ++t) gacc[t] = d4v{0, 0, 0, 0};
    // chunks [gcb, gcb + gn) of the rows (all of them in MODE 0)
    const int gcb = MODE == 1 ? split * nch / S.nsplit : 0;
    const int gn = (MODE == 1 ? (split + 1) * nch / S.nsplit : nch) - gcb;
    double pre[RING][PER];
#pragma unroll
    for (int u = 0; u < RING; ++u)
      if (u < gn) load_chunk(S.B, S.ldb, (gcb + u) * RCH, pre[u]);
    if (gn > 0) gstage(gbuf0, pre[0]);
    if (RING < gn) load_chunk(S.B, S.ldb, (gcb + RING) * RCH, pre[0]);
    __syncthreads();
    for (int c0 = 0; c0 < gn; c0 += RING) {
#pragma unroll
      for (int u = 0; u < RING; ++u) {
        const int c = c0 + u;
        if (c < gn) {  // uniform over the workgroup
          const int un = (u + 1) % RING;
          if (c + 1 < gn) {
            gstage((c & 1) ? gbuf0 : gbuf1, pre[un]);
            if (c + 1 + RING < gn) load_chunk(S.B, S.ldb, (gcb + c + 1 + RING) * RCH, pre[un]);
          }
          const double (*gch)[CP] = (c & 1) ? gbuf1 : gbuf0;
#pragma unroll
          for (int t = 0; t < C::TPW; ++t) {
            const int k = wave + NW * t;
            if (k < C::NGT) {  // uniform over the wave
              const int bi = int(C::GTILES >> (4 * k)) & 3, bj = int(C::GTILES >> (4 * k + 2)) & 3;
#pragma unroll
              for (int kk = 0; kk < RCH; kk += 4) {
                const double a = gch[16 * bi + (lane & 15)][kk + (lane >> 4)];
                const double b = gch[16 * bj + (lane & 15)][kk + (lane >> 4)];
                gacc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, gacc[t], 0, 0, 0);
              }
            }
          }
          __syncthreads();
        }
      }
    }
    double* gp = gpart + (size_t(bid) * gstride + split) * WC * WC;
#pragma unroll
    for (int t = 0; t < C::TPW; ++t) {
      const int k = wave + NW * t;
      if (k < C::NGT) {
        const int bi = int(C::GTILES >> (4 * k)) & 3, bj = int(C::GTILES >> (4 * k + 2)) & 3;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = 16 * bi + 4 * q + (lane >> 4), j = 16 * bj + (lane & 15);
          if (MODE == 1) gp[i * WC + j] = gacc[t][q];
          else G[i][j] = gacc[t][q];
        }
      }
    }
  }
  if (MODE == 1) {
    if (tid == 0 && S.info) atomicAdd(reinterpret_cast<unsigned*>(S.info) + 1, unsigned(wall_clock64() - t0));
    return;
  }
  __syncthreads();
  if (tid == 0) t1 = wall_clock64();
  double (*V)[GLD] = GA;
  if (MODE != 3) {
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
  const double tol2 = S.tol * S.tol;
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
    if (tid == 0) {
      if (MODE == 2) rotf[bid] = 0;
      if (S.info) {
        if (MODE == 0) atomicAdd(reinterpret_cast<unsigned*>(S.info) + 1, unsigned(t1 - t0));
        atomicAdd(reinterpret_cast<unsigned*>(S.info) + 4, 1u);
      }
    }
    return;
  }

  inner_sweep<WB_>(GA, GB, tol2, tid, &vrot);
  chunk = GB;
  if (tid == 0) {
    t2 = wall_clock64();
    if (S.info) {
      if (MODE == 0) atomicAdd(reinterpret_cast<unsigned*>(S.info) + 1, unsigned(t1 - t0));
      atomicAdd(reinterpret_cast<unsigned*>(S.info) + 2, unsigned(t2 - t1));
      atomicAdd(reinterpret_cast<unsigned*>(S.info) + 4, 1u);
    }
    if (MODE == 2) rotf[bid] = vrot;
  }
  if (!vrot) return;
  if (tid == 0) {
    flags[(sweep & 1) * njobs + lo] = 1;
    if (S.info) atomicAdd(reinterpret_cast<unsigned*>(S.info) + 5, 1u);
  }
  if (MODE == 2) {  // V to the buffer the apply launch reads
    double* vb = vbuf + size_t(bid) * WC * WC;
    for (int e = tid; e < WC * WC; e += JT) vb[e] = V[e / WC][e % WC];
    return;
  }
  } else {  // MODE 3
    if (rotf[bid] == 0) return;
    if (tid == 0) t2 = wall_clock64();
  }

  // ---- apply V to the visited columns of B and J: C <- C V (wave w: row tile w / NT,
  // column tile w % NT); the chunks of B and then of J form one sequence, so the
  // loads of J's first chunks overlap B's last MFMAs
  {
    // chunks [acb, acb + tot) of the sequence (all 2 nch of them in MODE 0)
    const int acb = MODE == 3 ? split * (2 * nch) / S.nsplit : 0;
    const int tot = (MODE == 3 ? (split + 1) * (2 * nch) / S.nsplit : 2 * nch) - acb;
    auto mat = [&](int c) { return acb + c < nch ? S.B : S.J; };
    auto mld = [&](int c) { return acb + c < nch ? S.ldb : S.ldj; };
    auto row0 = [&](int c) { return ((acb + c) % nch) * RCH; };
    double pre[RING][PER];
#pragma unroll
    for (int u = 0; u < RING; ++u)
      if (u < tot) load_chunk(mat(u), mld(u), row0(u), pre[u]);
    // wave w: row tile bi = w / NT, column tile bj = w % NT of each chunk
    const int bi = wave / NT, bj = wave % NT;
    // this wave's MFMA B operands (column tile bj of V) for every k, read once;
    // V's buffer (GA) then becomes the second chunk buffer beside GB
    double vb[WC / 4];
    const double* vg = vbuf + size_t(bid) * WC * WC;
#pragma unroll
    for (int kk = 0; kk < WC; kk += 4)
      vb[kk / 4] = MODE == 3 ? vg[(kk + (lane >> 4)) * WC + 16 * bj + (lane & 15)]
                             : V[kk + (lane >> 4)][16 * bj + (lane & 15)];
    __syncthreads();
    double(*abuf0)[GLD] = GB;
    double(*abuf1)[GLD] = GA;
    if (tot > 0) stage_chunk(abuf0, pre[0]);
    if (RING < tot) load_chunk(mat(RING), mld(RING), row0(RING), pre[0]);
    __syncthreads();
    const int gc = cols[16 * bj + (lane & 15)];
    for (int c0 = 0; c0 < tot; c0 += RING) {
#pragma unroll
      for (int u = 0; u < RING; ++u) {
        const int c = c0 + u;
        if (c < tot) {  // uniform over the workgroup
          const int un = (u + 1) % RING;
          if (c + 1 < tot) {
            stage_chunk((c & 1) ? abuf0 : abuf1, pre[un]);
            const int cn = c + 1 + RING;
            if (cn < tot) load_chunk(mat(cn), mld(cn), row0(cn), pre[un]);
          }
          const double(*ch)[GLD] = (c & 1) ? abuf1 : abuf0;
          double* M = mat(c);
          const int64_t ld = mld(c);
          const int r0 = row0(c);
          d4v acc = {0, 0, 0, 0};
#pragma unroll
          for (int kk = 0; kk < WC; kk += 4) {
            const double av = ch[16 * bi + (lane & 15)][kk + (lane >> 4)];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, vb[kk / 4], acc, 0, 0, 0);
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int row = r0 + 16 * bi + 4 * q + (lane >> 4);
            if (gc >= 0 && row < p) M[row + int64_t(gc) * ld] = acc[q];
          }
          __syncthreads();
        }
      }
    }
  }
  if (tid == 0 && S.info) atomicAdd(reinterpret_cast<unsigned*>(S.info) + 3, unsigned(wall_clock64() - t2));
}

// singular values: column norms of B, CN_COLS columns per workgroup (blockIdx.y),
// four per wave at once: lane-strided partial sums, then a butterfly
constexpr int CN_COLS = 16;
// Also clears the sort's change flag and, when given, the sweep's rotation flag
// of each matrix (both read only by later launches on the stream).
__global__ __launch_bounds__(256) void column_norms_kernel(const SVDJob* __restrict__ jobs, int32_t* changed,
                                                           int32_t* rot_flags) {
  if (blockIdx.y == 0 && threadIdx.x == 0) {
    if (changed) changed[blockIdx.x] = 0;
    if (rot_flags) rot_flags[blockIdx.x] = 0;
  }
  const SVDJob S = jobs[blockIdx.x];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c0 = blockIdx.y * CN_COLS + 4 * wave;
  if (c0 >= S.p) return;
  const double* x[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) x[q] = S.B + int64_t(min(c0 + q, S.p - 1)) * S.ldb;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = lane; i < S.p; i += 64)
#pragma unroll
    for (int q = 0; q < 4; ++q) s[q] += x[q][i] * x[q][i];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s[q] += __shfl_xor(s[q], off);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (c0 + q < S.p) S.sv[c0 + q] = sqrt(s[q]);
}

// Column order of a Jacobi matrix before a sweep: the stable order of decreasing
// norm (equal norms keep their order), as the rank of each column, perm[rank] = column.
// A NaN norm ranks below every number, so the ranks always form a permutation.
// changed[q] is set when matrix q's order is not the identity.
struct PermJob {
  double *B, *J, *tB, *tJ;  // the matrix and its rotations; p x p scratch for both
  int64_t ldb, ldj;
  const double* nrm;
  int32_t* perm;
  int32_t p;
};

__device__ __forceinline__ double sort_key(double v) { return isnan(v) ? -INFINITY : v; }

__global__ __launch_bounds__(256) void sort_perm_kernel(const PermJob* __restrict__ jobs, int32_t* __restrict__ changed) {
  const PermJob P = jobs[blockIdx.x];
  const int i = blockIdx.y * 256 + threadIdx.x;
  __shared__ double v[256];
  const double vi = i < P.p ? sort_key(P.nrm[i]) : 0.0;
  int rank = 0;
  for (int j0 = 0; j0 < P.p; j0 += 256) {
    __syncthreads();
    if (j0 + int(threadIdx.x) < P.p) v[threadIdx.x] = sort_key(P.nrm[j0 + threadIdx.x]);
    __syncthreads();
    const int n = min(256, P.p - j0);
    for (int t = 0; t < n; ++t) {
      const double vj = v[t];
      rank += (vj > vi || (vj == vi && j0 + t < i)) ? 1 : 0;
    }
  }
  if (i < P.p) {
    P.perm[rank] = i;
    if (rank != i) atomicOr(changed + blockIdx.x, 1);
  }
}

// tB(:, c) = B(:, perm[c]) and tJ(:, c) = J(:, perm[c]) for the matrices whose order changed
__global__ __launch_bounds__(256) void permute_gather_kernel(const PermJob* __restrict__ jobs,
                                                             const int32_t* __restrict__ changed) {
  if (!changed[blockIdx.y]) return;
  const PermJob P = jobs[blockIdx.y];
  const int64_t tot = int64_t(P.p) * P.p;
  for (int64_t e = blockIdx.x * int64_t(256) + threadIdx.x; e < tot; e += int64_t(gridDim.x) * 256) {
    const int64_t r = e % P.p, c = e / P.p;
    const int64_t src = min(max(P.perm[c], 0), P.p - 1);
    P.tB[e] = P.B[r + P.ldb * src];
    P.tJ[e] = P.J[r + P.ldj * src];
  }
}

__global__ __launch_bounds__(256) void permute_back_kernel(const PermJob* __restrict__ jobs,
                                                           const int32_t* __restrict__ changed) {
  if (!changed[blockIdx.y]) return;
  const PermJob P = jobs[blockIdx.y];
  const int64_t tot = int64_t(P.p) * P.p;
  for (int64_t e = blockIdx.x * int64_t(256) + threadIdx.x; e < tot; e += int64_t(gridDim.x) * 256) {
    const int64_t r = e % P.p, c = e / P.p;
    P.B[r + P.ldb * c] = P.tB[e];
    P.J[r + P.ldj * c] = P.tJ[e];
  }
}

// pseudo-inverse weights: one workgroup per matrix, the maximum singular value by
// a workgroup reduction, then column c of BJ scaled by 1/sv_c^2 or zeroed
__global__ __launch_bounds__(256) void pinv_weights_kernel(const PinvWeights* __restrict__ jobs) {
  const PinvWeights J = jobs[blockIdx.x];
  __shared__ double red[4];
  double mx = 0.0;
  for (int c = threadIdx.x; c < J.p; c += 256) mx = fmax(mx, J.sv[c]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  const double cut = J.eps * mx;
  const int64_t tot = int64_t(J.p) * J.p;
  for (int64_t e = threadIdx.x; e < tot; e += 256) {
    const int64_t i = e % J.p, c = e / J.p;
    const double s = J.sv[c];
    J.BJ[i + c * J.ld] *= s > cut ? 1.0 / (s * s) : 0.0;
  }
}

}  // namespace

void pinv_weights_batched(Ctx& ctx, const std::vector<PinvWeights>& jobs) {
  if (jobs.empty()) return;
  const PinvWeights* d = ctx.upload(jobs);
  TTR_LAUNCH(ctx, pinv_weights_kernel, dim3(unsigned(jobs.size())), dim3(256), 0, d);
  TTR_HIP(hipGetLastError());
}

template <int WB>
void jacobi_run(Ctx& ctx, std::vector<SVDJob>& jobs) {
  WsScope scope(ctx);  // descriptors, flags, partial Grams and sort buffers are needed only here
  std::vector<SVDJob> live;
  for (auto& j : jobs)
    if (j.p > 0) live.push_back(j);
  if (live.empty()) return;
  const int nj = int(live.size());
  std::vector<JacJob> jj(nj);
  std::vector<Fill2D> fills;
  int slots = 0, max_rounds = 1, max_split = 1;
  for (int q = 0; q < nj; ++q) {
    const SVDJob& s = live[q];
    JacJob& J = jj[q];
    J.B = s.B; J.ldb = s.ldb; J.J = s.J; J.ldj = s.ldj; J.info = s.info; J.p = s.p;
    J.tol = double(std::max(s.p, 16)) * 2.220446049250313e-16 * ctx.jacobi_tol_scale;
    J.nblk = (s.p + WB - 1) / WB;
    J.pp = J.nblk + (J.nblk & 1);
    J.rounds = J.nblk == 1 ? 1 : J.pp - 1;
    // row ranges of the split launches: about two 32-row chunks of Gram rows per
    // workgroup, by the matrix's size alone (a matrix's result does not depend on
    // the batch it comes in)
    J.nsplit = std::max(1, std::min(JAC_MAX_SPLIT, ((s.p + RCH - 1) / RCH + 1) / 2));
    max_split = std::max(max_split, J.nsplit);
    J.slot0 = slots;
    slots += J.nblk == 1 ? 1 : J.pp / 2;
    max_rounds = std::max(max_rounds, J.rounds);
    fills.push_back(Fill2D{s.J, s.ldj, s.p, s.p, 1.0});
  }
  const JacJob* dj = nullptr;
  const SVDJob* dl = nullptr;
  // split launches when some matrix has more than one row range: partial Grams,
  // V and the rotate flag of every visit slot
  const bool split = max_split > 1;
  constexpr int WC = Jac<WB>::WC;
  double* gpart = split ? ctx.ws.alloc_n<double>(size_t(slots) * max_split * WC * WC) : nullptr;
  double* vbuf = split ? ctx.ws.alloc_n<double>(size_t(slots) * WC * WC) : nullptr;
  int32_t* rotf = split ? ctx.ws.alloc_n<int32_t>(size_t(slots)) : nullptr;
  int32_t* flags = ctx.ws.alloc_n<int32_t>(size_t(2) * nj);
  const int max_sweeps = std::max(1, std::min(MAX_SWEEPS, ctx.jacobi_max_sweeps));
  // de Rijk's pivoting: before every sweep the columns of B (and, to keep
  // B = B_in J, of J) are sorted by decreasing norm, so each column block holds
  // columns of similar size.  On graded matrices (terms whose norms differ by
  // many orders of magnitude) the unsorted iteration needed more than 40 sweeps.
  std::vector<int64_t> off(size_t(nj) + 1, 0);
  int max_p = 1;
  for (int q = 0; q < nj; ++q) {
    off[q + 1] = off[q] + live[q].p;
    max_p = std::max(max_p, live[q].p);
  }
  const int cn_y = (max_p + CN_COLS - 1) / CN_COLS;  // column-norm workgroups per matrix
  const SVDJob* dn = nullptr;
  const PermJob* dperm = nullptr;
  int32_t* changed = nullptr;
  std::vector<SVDJob> nj_jobs;
  std::vector<PermJob> pj;
  if (ctx.jacobi_sort) {
    double* nrm = ctx.ws.alloc_n<double>(size_t(off[nj]));
    int32_t* perm = ctx.ws.alloc_n<int32_t>(size_t(off[nj]));
    changed = ctx.ws.alloc_n<int32_t>(size_t(nj));
    nj_jobs = live;
    pj.resize(static_cast<size_t>(nj));
    for (int q = 0; q < nj; ++q) {
      const int p = live[q].p;
      nj_jobs[q].sv = nrm + off[q];
      pj[q] = PermJob{live[q].B, live[q].J, ctx.ws.alloc_n<double>(size_t(p) * p), ctx.ws.alloc_n<double>(size_t(p) * p),
                      live[q].ldb, live[q].ldj, nrm + off[q], perm + off[q], p};
    }
  }
  {  // the run's descriptors in one upload: Jacobi jobs, SVD jobs, then the sort's jobs
    auto al = [](size_t x) { return (x + 15) & ~size_t(15); };
    const size_t o1 = al(jj.size() * sizeof(JacJob)), o2 = o1 + al(live.size() * sizeof(SVDJob)),
                 o3 = o2 + al(nj_jobs.size() * sizeof(SVDJob)), o4 = o3 + pj.size() * sizeof(PermJob);
    std::vector<char> pack(o4);
    std::memcpy(pack.data(), jj.data(), jj.size() * sizeof(JacJob));
    std::memcpy(pack.data() + o1, live.data(), live.size() * sizeof(SVDJob));
    if (!nj_jobs.empty()) std::memcpy(pack.data() + o2, nj_jobs.data(), nj_jobs.size() * sizeof(SVDJob));
    if (!pj.empty()) std::memcpy(pack.data() + o3, pj.data(), pj.size() * sizeof(PermJob));
    const char* d = ctx.upload(pack);
    dj = reinterpret_cast<const JacJob*>(d);
    dl = reinterpret_cast<const SVDJob*>(d + o1);
    if (!nj_jobs.empty()) {
      dn = reinterpret_cast<const SVDJob*>(d + o2);
      dperm = reinterpret_cast<const PermJob*>(d + o3);
    }
  }
  fill2d_batched(ctx, fills);
  zero_fill(ctx, flags, sizeof(int32_t) * 2 * nj);
  // the sort runs on the device (no host round trip per sweep)
  const unsigned perm_x = unsigned(std::min<int64_t>(64, (int64_t(max_p) * max_p + 255) / 256));
  auto sort_columns = [&](int32_t* rot_flags) {
    TTR_LAUNCH(ctx, column_norms_kernel, dim3(unsigned(nj), unsigned(cn_y)), dim3(256), 0, dn, changed, rot_flags);
    TTR_LAUNCH(ctx, sort_perm_kernel, dim3(unsigned(nj), unsigned((max_p + 255) / 256)), dim3(256), 0, dperm, changed);
    TTR_LAUNCH(ctx, permute_gather_kernel, dim3(perm_x, unsigned(nj)), dim3(256), 0, dperm,
               static_cast<const int32_t*>(changed));
    TTR_LAUNCH(ctx, permute_back_kernel, dim3(perm_x, unsigned(nj)), dim3(256), 0, dperm,
               static_cast<const int32_t*>(changed));
    TTR_HIP(hipGetLastError());
  };
  // The flags of sweep s are copied to pinned memory behind its launches and read
  // while sweep s + 1 runs: the host never waits for the sweep it has just queued.
  // The sweep after the last rotating one leaves every matrix unchanged (its
  // workgroups return at once, its columns are already in order).
  int32_t* hfl = static_cast<int32_t*>(ctx.pin.alloc(sizeof(int32_t) * 2 * size_t(nj)));
  struct Events {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Events() {
      for (auto x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } ev;
  for (auto& x : ev.e) TTR_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
  std::vector<int> rotating;  // matrices still rotating after each checked sweep
  int last_checked = -1;
  auto still_rotating = [&](int sw) {  // matrices that rotated in sweep sw (copied flags)
    TTR_HIP(hipEventSynchronize(ev.e[sw & 1]));
    const int32_t* f = hfl + size_t(sw & 1) * nj;
    int cnt = 0;
    for (int q = 0; q < nj; ++q) cnt += f[q] != 0;
    rotating.push_back(cnt);
    last_checked = sw;
    return cnt;
  };
  bool converged = false;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    int32_t* cur = flags + size_t(sweep & 1) * nj;
    // the flags of sweep - 2 were copied out; they start at zero again
    if (ctx.jacobi_sort) sort_columns(sweep >= 2 ? cur : nullptr);
    else if (sweep >= 2) zero_fill(ctx, cur, sizeof(int32_t) * nj);
    for (int r = 0; r < max_rounds; ++r) {
      // timed per launch; the work depends on the sweeps taken, so the caller
      // adds it from the counters
      ProfScope ps(ctx, "jacobi_svd");
      if (!split) {
        TTR_LAUNCH(ctx, (jacobi_round_kernel<WB, 0>), dim3(unsigned(slots)), dim3(Jac<WB>::JT), 0, dj, nj, r, sweep,
                   flags, gpart, max_split, vbuf, rotf);
      } else {
        TTR_LAUNCH(ctx, (jacobi_round_kernel<WB, 1>), dim3(unsigned(slots), unsigned(max_split)), dim3(Jac<WB>::JT),
                   0, dj, nj, r, sweep, flags, gpart, max_split, vbuf, rotf);
        TTR_LAUNCH(ctx, (jacobi_round_kernel<WB, 2>), dim3(unsigned(slots)), dim3(Jac<WB>::JT), 0, dj, nj, r, sweep,
                   flags, gpart, max_split, vbuf, rotf);
        TTR_LAUNCH(ctx, (jacobi_round_kernel<WB, 3>), dim3(unsigned(slots), unsigned(max_split)), dim3(Jac<WB>::JT),
                   0, dj, nj, r, sweep, flags, gpart, max_split, vbuf, rotf);
      }
      TTR_HIP(hipGetLastError());
    }
    // convergence checked from the third sweep on (no matrix converges earlier in
    // practice) and after the last sweep allowed, one sweep late
    if (sweep >= 2 || sweep == max_sweeps - 1) {
      TTR_HIP(hipMemcpyAsync(hfl + size_t(sweep & 1) * nj, cur, sizeof(int32_t) * nj, hipMemcpyDeviceToHost,
                             ctx.stream));
      TTR_HIP(hipEventRecord(ev.e[sweep & 1], ctx.stream));
    }
    if (sweep - 1 >= 2 && still_rotating(sweep - 1) == 0) {
      converged = true;
      break;
    }
  }
  if (!converged && last_checked != max_sweeps - 1) converged = still_rotating(max_sweeps - 1) == 0;
  // a matrix that still rotated in the last sweep has no converged SVD: report it
  // (the reference's LAPACK driver either converges or throws, src/tt_rounding.jl:191-195)
  if (!converged) {
    std::string msg = "Jacobi SVD did not converge in " + std::to_string(max_sweeps) +
                      " sweeps (matrices still rotating after each sweep from the third:";
    for (int c : rotating) msg += " " + std::to_string(c);
    msg += "; sizes of the unconverged ones:";
    int shown = 0;
    const int32_t* f = hfl + size_t(last_checked & 1) * nj;
    for (int q = 0; q < nj && shown < 16; ++q)
      if (f[q]) {
        msg += " " + std::to_string(live[q].p);
        ++shown;
      }
    throw Error(ST_NUMERICAL_FAILURE, msg + ")");
  }
  TTR_LAUNCH(ctx, column_norms_kernel, dim3(unsigned(nj), unsigned(cn_y)), dim3(256), 0, dl, nullptr, nullptr);
  TTR_HIP(hipGetLastError());
}


// Column blocks of 32 (64-column visits, 512 threads).  TTR_JACOBI_WB=16
// selects 32-column visits: the same number of sweeps and half the inner
// rotation work per sweep, but four times the visits, whose Gram and apply
// phases read every column twice as often (measured 11% slower on the square
// factors of the rounding sweeps).
int jacobi_svd_batched(Ctx& ctx, std::vector<SVDJob>& jobs) {
  static const int wb = [] {
    const char* e = std::getenv("TTR_JACOBI_WB");
    return (e && std::atoi(e) == 16) ? 16 : 32;
  }();
  if (wb == 32) jacobi_run<32>(ctx, jobs);
  else jacobi_run<16>(ctx, jobs);
  return 2 * wb;
}

}  // namespace ttr
