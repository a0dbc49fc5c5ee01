// Batched FP64 Householder QR with explicit thin Q.
//
// Two-level blocked right-looking Householder in compact WY form
// (I - V T V^T), following the conventions of LAPACK dgeqrf/dlarfg/dlarft/
// dorgqr so that Q agrees with the reference's LAPACK factorisations column by
// column:
//   beta = -sign(alpha) * ||(alpha, x)||,  tau = (beta - alpha)/beta,
//   v = (1, x/(alpha - beta)),  tau = 0 when x == 0.
// Columns are processed in blocks of 64.  Up to 3072 rows one workgroup per
// matrix factors the whole block in one launch (panels of 16 in registers, each
// applied on chip to the rest of the block); taller matrices take panels of 32
// factorised by a streaming kernel and applied to the rest of the block by
// grouped GEMMs.  The block's reflectors are then merged into one 64-wide
// compact WY factor (T from V^T V and the panels' T's), which updates the
// trailing matrix and, backwards, forms Q, as grouped FP64 MFMA GEMMs over all
// matrices of the batch, so the trailing matrix is streamed once per 64 columns.
#include "ttr_internal.h"
#include "qr_step.h"

namespace ttr {

namespace {

constexpr int NBP = 32;     // panel width
constexpr int PT = 256;     // panel kernel threads
constexpr int PW = PT / 64; // waves

struct PanelJob {
  double* A;
  int64_t lda;
  int32_t m, c0, nb;
  double* tau;  // tau for columns c0..c0+nb-1 (indexed by column)
  double* T;    // NBP x NBP, ld NBP, upper triangular
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__global__ __launch_bounds__(PT) void qr_panel_kernel(const PanelJob* __restrict__ jobs) {
  const PanelJob J = jobs[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ double red[PW][NBP + 1];
  __shared__ double dsum[NBP];
  __shared__ double Ts[NBP][NBP + 1];
  __shared__ double sh_alpha, sh_sigma;

  for (int idx = tid; idx < NBP * (NBP + 1); idx += PT) (&Ts[0][0])[idx] = 0.0;

  double* A = J.A;
  const int64_t lda = J.lda;
  const int m = J.m;
  for (int jj = 0; jj < J.nb; ++jj) {
    const int j = J.c0 + jj;
    double* colj = A + int64_t(j) * lda;
    // --- norm of the sub-column
    double part = 0.0;
    for (int i = j + 1 + tid; i < m; i += PT) {
      const double x = colj[i];
      part += x * x;
    }
    part = wave_sum(part);
    if (lane == 0) red[wave][0] = part;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < PW; ++w) s += red[w][0];
      sh_sigma = s;
      sh_alpha = colj[j];
    }
    __syncthreads();
    const double sigma = sh_sigma, alpha = sh_alpha;
    double tau = 0.0, beta = alpha, scal = 1.0;
    if (sigma != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    if (sigma != 0.0)
      for (int i = j + 1 + tid; i < m; i += PT) colj[i] *= scal;
    __syncthreads();
    if (tid == 0) {
      colj[j] = beta;
      J.tau[j] = tau;
    }
    // --- d_c = A[j,c] + sum_{i>j} v_i A[i,c] for every other panel column c:
    //     c > j: w_c for the reflector application; c < j: V^T v_j for T
    double acc[NBP];
#pragma unroll
    for (int c = 0; c < NBP; ++c) acc[c] = 0.0;
    for (int i = j + 1 + tid; i < m; i += PT) {
      const double v = colj[i];
#pragma unroll
      for (int c = 0; c < NBP; ++c)
        if (c < J.nb && c != jj) acc[c] += v * A[i + int64_t(J.c0 + c) * lda];
    }
#pragma unroll
    for (int c = 0; c < NBP; ++c) {
      const double s = wave_sum(acc[c]);
      if (lane == 0) red[wave][c] = s;
    }
    __syncthreads();
    if (tid < J.nb && tid != jj) {
      double s = A[j + int64_t(J.c0 + tid) * lda];
      for (int w = 0; w < PW; ++w) s += red[w][tid];
      dsum[tid] = s;
    }
    __syncthreads();
    // --- apply H_j to the remaining panel columns (rows j..m-1)
    if (tau != 0.0) {
      for (int c = jj + 1; c < J.nb; ++c) {
        const double w = tau * dsum[c];
        double* colc = A + int64_t(J.c0 + c) * lda;
        for (int i = j + tid; i < m; i += PT) {
          const double v = (i == j) ? 1.0 : colj[i];
          colc[i] -= v * w;
        }
      }
    }
    // --- T(:, jj): T(jj,jj) = tau, T(0:jj, jj) = -tau * T(0:jj,0:jj) z
    if (tid < jj) {
      double s = 0.0;
      for (int q = tid; q < jj; ++q) s += Ts[tid][q] * dsum[q];
      Ts[tid][jj] = -tau * s;
    }
    if (tid == 0) Ts[jj][jj] = tau;
    __syncthreads();
  }
  for (int idx = tid; idx < NBP * NBP; idx += PT) {
    const int r = idx % NBP, c = idx / NBP;
    J.T[r + NBP * c] = (r < J.nb && c < J.nb) ? Ts[r][c] : 0.0;
  }
}


// Register-resident block factorisation, one launch per block of NBO = 64
// columns: one 512-thread workgroup per matrix factors the block's panels of
// NBR = 16 columns in registers (thread t holds panel rows t, t+512, ...; the
// column steps of qr_step.h) and applies each panel to the block's columns right
// of it on chip, so the in-block updates (K = 16) are no grid-wide GEMMs.  Each
// panel's explicit V (unit diagonal, zeros above) is written to the block's V
// buffer, which the update reads back (L2) and the trailing update and the
// formation of Q read afterwards.  The update splits the rows over the 8 waves
// in multiples of 16 (v_mfma_f64_16x16x4):
//   W = V^T C   (16 x n2: per-wave partial sums, added in LDS in wave order),
//   W2 = T^T W, C -= V W2   (each wave on its own rows).
constexpr int NBR = qrstep::SNB;
constexpr int RT = 512;
constexpr int BW = RT / 64;  // waves
constexpr int NBO = 64;
constexpr int BNS = (NBO - NBR) / NBR;  // slabs of 16 columns right of a panel (at most 3)
constexpr int BU = 4;                   // k-steps (4 rows each) per batch of loads

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef const double __attribute__((address_space(1)))* gcd_t;
typedef double __attribute__((address_space(1)))* gd_t;

struct BlockJob {
  double* A;
  int64_t lda;
  int32_t m, c0, nbo;  // block columns c0 .. c0 + nbo - 1 (nbo <= NBO)
  double* tau;         // indexed by column
  double* T;           // T of the panel at columns 16 q .. at T + q * NBP * NBP, ld NBP
  double* VX;          // the block's explicit V: (m - c0) x nbo, ld vld
  int64_t vld;
  int32_t vz;          // rows above VX (of the same columns) set to zero: 0, or 64 for the
                       // second block of a 128-column pair
};

constexpr int QR_ARG_JOBS = 4;  // jobs of a QR block or T-merge launch passed in its arguments

template <int R, bool INL>
__global__ __launch_bounds__(RT) void qr_block_kernel(const BlockJob* __restrict__ jobs,
                                                      ArgJobs<BlockJob, QR_ARG_JOBS, INL> aj) {
  if constexpr (INL) jobs = aj.v;
  const BlockJob J = jobs[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  __shared__ qrstep::PanelShared16<RT> sh;
  __shared__ double Wred[BW][BNS][NBR * NBR];
  __shared__ double Ws[BNS][NBR * NBR];
  double* A = J.A;
  const int64_t lda = J.lda;
  const int m = J.m, nbo = J.nbo;
  const int64_t mb = J.vld;  // the ld of VX
  for (int e = tid; e < J.vz * nbo; e += RT) J.VX[e % J.vz - J.vz + mb * (e / J.vz)] = 0.0;
  for (int p = 0; NBR * p < nbo; ++p) {
    const int pc = J.c0 + NBR * p, nb = min(NBR, nbo - NBR * p), m2 = m - pc;
    double* A0 = A + pc + lda * pc;
    double* V0 = J.VX + NBR * p + mb * (NBR * p);  // the panel's V: row i, column c at V0[i + mb c]
    double P[R][NBR];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = tid + RT * q;
#pragma unroll
      for (int c = 0; c < NBR; ++c) P[q][c] = (i < m2 && c < nb) ? A0[i + lda * c] : 0.0;
    }
    for (int idx = tid; idx < NBR * qrstep::LDV; idx += RT) sh.Ts[idx] = 0.0;
    __syncthreads();
    qrstep::PanelStep16<R, RT, 0>::run(P, sh, J.tau + pc, nb, m2, tid, lane, wave);
    // R and V back to A; explicit V to VX (zero in the block's rows above the panel)
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = tid + RT * q;
      if (i < m2) {
#pragma unroll
        for (int c = 0; c < NBR; ++c)
          if (c < nb) {
            A0[i + lda * c] = P[q][c];
            V0[i + mb * c] = i > c ? P[q][c] : (i == c ? 1.0 : 0.0);
          }
      }
    }
    for (int e = tid; e < NBR * p * nb; e += RT) J.VX[e % (NBR * p) + mb * (NBR * p + e / (NBR * p))] = 0.0;
    __syncthreads();
    {
      double* Tp = J.T + size_t(pc / NBR) * NBP * NBP;
      for (int idx = tid; idx < NBP * NBP; idx += RT) {
        const int r = idx % NBP, c = idx / NBP;
        Tp[idx] = (r < nb && c < nb && r < NBR && c < NBR) ? sh.Ts[r * qrstep::LDV + c] : 0.0;
      }
    }
    const int n2 = nbo - NBR * p - nb;  // block columns right of the panel
    if (n2 > 0) {
      const int ns = (n2 + NBR - 1) / NBR;
      // global-address-space views: global (not flat) loads and stores, which LDS
      // waits do not wait for
      const gd_t C = (gd_t)(A0 + lda * nb);
      const gcd_t Vg = (gcd_t)V0;
      const int rpw = (m2 + NBR * BW - 1) / (NBR * BW) * NBR;
      const int r0 = min(wave * rpw, m2), r1 = min(r0 + rpw, m2);
      // W = V^T C over this wave's rows, batches of BU k-steps, the next batch's
      // loads issued before the current batch's MFMAs
      d4_t w[BNS];
#pragma unroll
      for (int s = 0; s < BNS; ++s) w[s] = d4_t{0.0, 0.0, 0.0, 0.0};
      // rows past r1 load row r1 - 1 (always valid) and meet a zero V entry;
      // columns past n2 load column n2 - 1 and give W columns nobody stores
      auto load_w = [&](double (&a)[BU], double (&b)[BNS][BU], int ib) {
#pragma unroll
        for (int u = 0; u < BU; ++u) {
          const int i = ib + 4 * u + lk;
          const int ic = min(i, r1 - 1);
          const double v = Vg[ic + mb * li];
          a[u] = i < r1 ? v : 0.0;
#pragma unroll
          for (int s = 0; s < BNS; ++s) b[s][u] = C[ic + lda * min(NBR * s + li, n2 - 1)];
        }
      };
      auto mma_w = [&](const double (&a)[BU], const double (&b)[BNS][BU]) {
#pragma unroll
        for (int u = 0; u < BU; ++u)
#pragma unroll
          for (int s = 0; s < BNS; ++s)
            if (s < ns) w[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[s][u], w[s], 0, 0, 0);
      };
      if (r0 < r1) {
        double a0[BU], b0[BNS][BU], a1[BU], b1[BNS][BU];
        load_w(a0, b0, r0);
        for (int ib = r0; ib < r1; ib += 8 * BU) {
          load_w(a1, b1, ib + 4 * BU);
          mma_w(a0, b0);
          if (ib + 4 * BU >= r1) break;
          load_w(a0, b0, ib + 8 * BU);
          mma_w(a1, b1);
        }
      }
      // register r of a 16 x 16 MFMA result holds row 4 r + lane / 16, column lane % 16
#pragma unroll
      for (int s = 0; s < BNS; ++s)
        if (s < ns)
#pragma unroll
          for (int r = 0; r < 4; ++r) Wred[wave][s][(4 * r + lk) * NBR + li] = w[s][r];
      __syncthreads();
      for (int e = tid; e < ns * NBR * NBR; e += RT) {
        const int s = e / (NBR * NBR), o = e % (NBR * NBR);
        double v = 0.0;
#pragma unroll
        for (int ww = 0; ww < BW; ++ww) v += Wred[ww][s][o];
        Ws[s][o] = v;
      }
      __syncthreads();
      // W2 = T^T W (every wave, for its own update)
      d4_t w2[BNS];
#pragma unroll
      for (int s = 0; s < BNS; ++s) {
        w2[s] = d4_t{0.0, 0.0, 0.0, 0.0};
        if (s < ns) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int k = 4 * r + lk;
            w2[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(sh.Ts[k * qrstep::LDV + li], Ws[s][k * NBR + li], w2[s], 0, 0,
                                                         0);
          }
        }
      }
      // C -= V W2 on this wave's rows, 16 at a time, the next 16 loaded before the
      // current ones are updated (clamped rows as above: loaded, never stored)
      auto load_u = [&](double (&a)[4], d4_t (&c)[BNS], int i0) {
        const int ia = min(i0 + li, r1 - 1);
#pragma unroll
        for (int r = 0; r < 4; ++r) a[r] = -Vg[ia + mb * (4 * r + lk)];
#pragma unroll
        for (int s = 0; s < BNS; ++s) {
          const gcd_t Cs = C + lda * min(NBR * s + li, n2 - 1);
#pragma unroll
          for (int r = 0; r < 4; ++r) c[s][r] = Cs[min(i0 + 4 * r + lk, r1 - 1)];
        }
      };
      auto upd = [&](const double (&a)[4], d4_t (&c)[BNS], int i0) {
#pragma unroll
        for (int s = 0; s < BNS; ++s)
          if (s < ns)
#pragma unroll
            for (int r = 0; r < 4; ++r) c[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], w2[s][r], c[s], 0, 0, 0);
#pragma unroll
        for (int s = 0; s < BNS; ++s) {
          const int j = NBR * s + li;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = i0 + 4 * r + lk;
            if (i < r1 && j < n2) C[i + lda * j] = c[s][r];
          }
        }
      };
      if (r0 < r1) {
        double aa[4], ab[4];
        d4_t ca[BNS], cb[BNS];
        load_u(aa, ca, r0);
        for (int i0 = r0; i0 < r1; i0 += 2 * NBR) {
          load_u(ab, cb, i0 + NBR);
          upd(aa, ca, i0);
          if (i0 + NBR >= r1) break;
          load_u(aa, ca, i0 + 2 * NBR);
          upd(ab, cb, i0 + NBR);
        }
      }
    }
    __syncthreads();
  }
}

// T of a block of nb <= 64 reflectors (dlarft, forward, columnwise) assembled from
// the T's its panels already have: with the block split into panels 1..q, the
// columns of panel q are T(0:c0, q) = -T(0:c0, 0:c0) G(0:c0, q) T_qq (G = V^T V of
// the block), T_qq the panel's own T.  One 256-thread workgroup per matrix, two
// small products per panel instead of one sequential step per column.
struct TMergeJob {
  const double* G;   // 64 x 64, ld ldg
  const double* Tp;  // panel q's T at Tp + q * NBP * NBP, ld NBP
  double* T;         // 64 x 64, ld ldt
  int32_t nb, w;     // columns of the block, panel width (16 or 32)
  int32_t ldg, ldt;
  int32_t zero_left; // also zero this many 64-column blocks left of T (the rows of a group's T below the diagonal)
};

template <bool INL>
__global__ __launch_bounds__(256) void merge_t_kernel(const TMergeJob* __restrict__ jobs,
                                                      ArgJobs<TMergeJob, QR_ARG_JOBS, INL> aj) {
  if constexpr (INL) jobs = aj.v;
  const TMergeJob J = jobs[blockIdx.x];
  const int t = threadIdx.x, nb = J.nb, w = J.w;
  __shared__ double Ts[NBO][NBO + 1];
  __shared__ double Gs[NBO][NBO + 1];
  __shared__ double X[NBO - 16][NBP + 1];
  for (int e = t; e < NBO * NBO; e += 256) {
    const int r = e % NBO, c = e / NBO;
    Gs[r][c] = (r < nb && c < nb) ? J.G[r + int64_t(J.ldg) * c] : 0.0;
    const int q = c / w;
    Ts[r][c] = (r < nb && c < nb && r / w == q) ? J.Tp[size_t(q) * NBP * NBP + (r - q * w) + NBP * (c - q * w)] : 0.0;
  }
  __syncthreads();
  for (int c0 = w; c0 < nb; c0 += w) {
    const int wq = min(w, nb - c0);
    // X (c0 x wq) = G(0:c0, c0:c0+wq) T_qq (T_qq upper triangular)
    for (int e = t; e < c0 * wq; e += 256) {
      const int r = e % c0, c = e / c0;
      double x = 0.0;
      for (int k = 0; k <= c; ++k) x += Gs[r][c0 + k] * Ts[c0 + k][c0 + c];
      X[r][c] = x;
    }
    __syncthreads();
    // T(0:c0, c0:c0+wq) = -T(0:c0, 0:c0) X (T upper triangular)
    for (int e = t; e < c0 * wq; e += 256) {
      const int r = e % c0, c = e / c0;
      double y = 0.0;
      for (int k = r; k < c0; ++k) y += Ts[r][k] * X[k][c];
      Ts[r][c0 + c] = -y;
    }
    __syncthreads();
  }
  for (int e = t; e < NBO * NBO; e += 256) J.T[e % NBO + int64_t(J.ldt) * (e / NBO)] = Ts[e % NBO][e / NBO];
  for (int e = t; e < J.zero_left * NBO * NBO; e += 256)
    J.T[e % NBO + int64_t(J.ldt) * (e / NBO - J.zero_left * NBO)] = 0.0;
}

void launch_panel(Ctx& ctx, const PanelJob* d, int njobs) {
  TTR_LAUNCH(ctx, qr_panel_kernel, dim3(unsigned(njobs)), dim3(PT), 0, d);
  TTR_HIP(hipGetLastError());
}

template <bool INL>
void launch_block_t(Ctx& ctx, const BlockJob* d, const ArgJobs<BlockJob, QR_ARG_JOBS, INL>& a, int njobs,
                    int max_rows) {
  const int R = (max_rows + RT - 1) / RT;
  const dim3 g{unsigned(njobs)}, b{unsigned(RT)};
  switch (R) {
    case 0:
    case 1: TTR_LAUNCH(ctx, (qr_block_kernel<1, INL>), g, b, 0, d, a); break;
    case 2: TTR_LAUNCH(ctx, (qr_block_kernel<2, INL>), g, b, 0, d, a); break;
    case 3: TTR_LAUNCH(ctx, (qr_block_kernel<3, INL>), g, b, 0, d, a); break;
    case 4: TTR_LAUNCH(ctx, (qr_block_kernel<4, INL>), g, b, 0, d, a); break;
    case 5: TTR_LAUNCH(ctx, (qr_block_kernel<5, INL>), g, b, 0, d, a); break;
    default: TTR_LAUNCH(ctx, (qr_block_kernel<6, INL>), g, b, 0, d, a); break;
  }
  TTR_HIP(hipGetLastError());
}

// one launch of qr_block_kernel over the jobs (timed as "qr_block")
void launch_block(Ctx& ctx, const std::vector<BlockJob>& bj, int max_rows, double fl, double by) {
  if (arg_jobs_fit<QR_ARG_JOBS>(ctx, bj)) {
    ProfScope ps(ctx, "qr_block", fl, by);
    launch_block_t(ctx, nullptr, arg_jobs<QR_ARG_JOBS>(bj), int(bj.size()), max_rows);
    return;
  }
  const BlockJob* d = ctx.upload(bj);
  ProfScope ps(ctx, "qr_block", fl, by);
  launch_block_t(ctx, d, ArgJobs<BlockJob, QR_ARG_JOBS, false>{}, int(bj.size()), max_rows);
}

void launch_merge_t(Ctx& ctx, const std::vector<TMergeJob>& tb) {
  if (tb.empty()) return;
  const dim3 g{unsigned(tb.size())}, b{256u};
  if (arg_jobs_fit<QR_ARG_JOBS>(ctx, tb)) {
    TTR_LAUNCH(ctx, merge_t_kernel<true>, g, b, 0, nullptr, arg_jobs<QR_ARG_JOBS>(tb));
  } else {
    const TMergeJob* d = ctx.upload(tb);
    TTR_LAUNCH(ctx, merge_t_kernel<false>, g, b, 0, d, (ArgJobs<TMergeJob, QR_ARG_JOBS, false>{}));
  }
}

// Blocked QR of matrices up to 6 * RT rows with the 64-column blocks taken in
// groups of G: block h of group s is updated by the group's first h blocks,
// factored (its panels and in-block updates in one launch, its explicit V written
// into the group's V with zero rows above it), and the group's compact WY factor
// grows by it: T_s = [T_acc, -T_acc (V_acc^T V_h) T_h; 0, T_h].  The group's 64 G
// reflectors then update the trailing columns and later form Q with K = 64 G, so
// those products pass over the trailing matrix once per 64 G columns (the update
// A2 -= V W2 measured 18.6, 30.2 and 42.1 TF/s at K = 64, 128 and 256).  G = 2 by
// default (Ctx::qr_group).
void qr_groups(Ctx& ctx, std::vector<QRJob>& jobs, int G) {
  const int NS = NBO * G;  // columns of a group
  struct Aux {
    double *tau, *T, *TS, *G, *Y, *W, *W2, *VX;
    bool vx_all;
    int kq, ngrp;
  };
  std::vector<Aux> aux(jobs.size());
  int max_grp = 0;
  for (size_t q = 0; q < jobs.size(); ++q) {
    const QRJob& J = jobs[q];
    TTR_REQUIRE(J.m >= 0 && J.n >= 0 && J.lda >= std::max(1, J.m), ST_INVALID_ARG, "qr: bad shape");
    Aux& a = aux[q];
    a.kq = std::min(J.m, J.n);
    a.ngrp = (a.kq + NS - 1) / NS;
    max_grp = std::max(max_grp, a.ngrp);
    const int npanel = (a.kq + NBR - 1) / NBR;
    a.tau = ctx.ws.alloc_n<double>(std::max(1, J.n));
    a.T = ctx.ws.alloc_n<double>(size_t(std::max(1, npanel)) * NBP * NBP);
    a.TS = ctx.ws.alloc_n<double>(size_t(std::max(1, a.ngrp)) * NS * NS);
    a.G = ctx.ws.alloc_n<double>(size_t(NS) * NBO);
    a.Y = ctx.ws.alloc_n<double>(size_t(NS) * NBO);
    a.W = ctx.ws.alloc_n<double>(size_t(NS) * std::max(1, J.n));
    a.W2 = ctx.ws.alloc_n<double>(size_t(NS) * std::max(1, J.n));
    // group s: (m - NS s) x NS, ld m - NS s; all groups kept when Q is formed
    a.vx_all = J.Q != nullptr;
    a.VX = ctx.ws.alloc_n<double>(
        a.vx_all ? size_t(NS) * (size_t(a.ngrp) * J.m - size_t(NS / 2) * a.ngrp * std::max(0, a.ngrp - 1))
                 : size_t(std::max(1, J.m)) * NS);
  }
  auto blk_cols = [&](const Aux& a, int b) { return std::max(0, std::min(NBO, a.kq - b * NBO)); };
  auto vxs = [&](const QRJob& J, const Aux& a, int s) {
    return a.VX + (a.vx_all ? int64_t(NS) * (int64_t(s) * J.m - int64_t(NS / 2) * s * (s - 1)) : 0);
  };
  auto ts = [&](const Aux& a, int s) { return a.TS + size_t(s) * NS * NS; };
  auto grp_cols = [&](const Aux& a, int s) { return std::max(0, std::min(NS, a.kq - s * NS)); };
  GemmBatch gb;
  // block h of group s of every job that has it: panels and in-block updates on chip
  auto factor_block = [&](int s, int h) {
    std::vector<BlockJob> bj;
    double fl = 0.0, by = 0.0;
    int max_rows = 0;
    for (size_t q = 0; q < jobs.size(); ++q) {
      const QRJob& J = jobs[q];
      Aux& a = aux[q];
      const int b = G * s + h, nbo = blk_cols(a, b);
      if (nbo <= 0) continue;
      const int c0 = b * NBO, mb = J.m - c0;
      const int64_t ld = J.m - int64_t(NS) * s;
      bj.push_back(BlockJob{J.A, J.lda, J.m, c0, nbo, a.tau, a.T, vxs(J, a, s) + int64_t(h) * NBO * (1 + ld), ld,
                            h * NBO});
      max_rows = std::max(max_rows, mb);
      for (int c = 0; c < nbo; c += NBR) {
        const double m2 = J.m - (c0 + c), nb = std::min(NBR, nbo - c), n2 = nbo - c - nb;
        fl += 4.0 * m2 * nb * nb + 4.0 * m2 * nb * n2;  // panel (2 m nb^2) + T, W = V^T C and C -= V W2
      }
      by += 24.0 * double(mb) * nbo;  // the block read and written, V written
    }
    if (bj.empty()) return;
    launch_block(ctx, bj, max_rows, fl, by);
  };
  // C -= V (T^T (V^T C)) (trans) or C -= V (T (V^T C)) for every job: V the first
  // kv columns of group s's V, C columns [j0, j1) of the job's A or Q from row NS s
  auto apply = [&](int s, bool trans, bool on_q, const char* site, auto kv_of, auto cols_of) {
    for (int stage = 0; stage < 3; ++stage) {
      for (size_t q = 0; q < jobs.size(); ++q) {
        const QRJob& J = jobs[q];
        const Aux& a = aux[q];
        if (s >= a.ngrp || (on_q && !J.Q)) continue;
        const int kv = kv_of(a, s);
        int j0, j1;
        cols_of(J, a, s, j0, j1);
        const int r0 = NS * s, m2 = J.m - r0, n2 = j1 - j0;
        if (kv <= 0 || n2 <= 0) continue;
        const int64_t ld = J.m - int64_t(NS) * s, ldc = on_q ? J.ldq : J.lda;
        double* C = (on_q ? J.Q : J.A) + r0 + ldc * j0;
        const double* V = vxs(J, a, s);
        const double* T = ts(a, s);
        if (stage == 0)  // W = V^T C  (kv x n2)
          gb.add(kv, n2, m2, 1.0, V, ld, 1, C, 1, ldc, 0.0, a.W, 1, NS);
        else if (stage == 1)  // W2 = T^T W or T W  (kv x n2)
          gb.add(kv, n2, kv, 1.0, T, trans ? NS : 1, trans ? 1 : NS, a.W, 1, NS, 0.0, a.W2, 1, NS);
        else  // C -= V W2
          gb.add(m2, n2, kv, -1.0, V, 1, ld, a.W2, 1, NS, 1.0, C, 1, ldc);
      }
      gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, site);
    }
  };
  // ---- factorisation
  for (int s = 0; s < max_grp; ++s) {
    for (int h = 0; h < G; ++h) {
      // the group's first h blocks applied to block h's columns
      if (h > 0)
        apply(s, true, false, "qr_group", [&](const Aux& a, int ss) { return blk_cols(a, G * ss + h) > 0 ? h * NBO : 0; },
              [&](const QRJob&, const Aux& a, int ss, int& j0, int& j1) {
                j0 = (G * ss + h) * NBO;
                j1 = j0 + blk_cols(a, G * ss + h);
              });
      factor_block(s, h);
      // G(0:64h+nb, 0:nb) = V_s^T V_h: rows 0..64h-1 hold V_acc^T V_h, the rest V_h^T V_h
      bool any = false;
      for (size_t q = 0; q < jobs.size(); ++q) {
        const QRJob& J = jobs[q];
        const Aux& a = aux[q];
        const int nb = blk_cols(a, G * s + h);
        if (nb <= 0) continue;
        const int64_t ld = J.m - int64_t(NS) * s;
        const double* V = vxs(J, a, s);
        gb.add(h * NBO + nb, nb, int(ld), 1.0, V, ld, 1, V + int64_t(h) * NBO * ld, 1, ld, 0.0, a.G, 1, NS);
        any = true;
      }
      if (!any) break;
      gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "qr_vtv");
      // T_h from rows 64h.. of G (and zeros left of it in its rows), then
      // T(0:64h, 64h:) = -T_acc (V_acc^T V_h) T_h
      std::vector<TMergeJob> tb;
      for (size_t q = 0; q < jobs.size(); ++q) {
        const Aux& a = aux[q];
        const int nb = blk_cols(a, G * s + h);
        if (nb <= 0) continue;
        tb.push_back(TMergeJob{a.G + h * NBO, a.T + size_t((G * s + h) * NBO / NBR) * NBP * NBP,
                               ts(a, s) + int64_t(h) * NBO * (1 + NS), nb, NBR, NS, NS, h});
      }
      launch_merge_t(ctx, tb);
      if (h == 0) continue;
      for (size_t q = 0; q < jobs.size(); ++q) {
        const Aux& a = aux[q];
        const int nb = blk_cols(a, G * s + h);
        if (nb <= 0) continue;
        gb.add(h * NBO, nb, nb, 1.0, a.G, 1, NS, ts(a, s) + int64_t(h) * NBO * (1 + NS), 1, NS, 0.0, a.Y, 1, NS);
      }
      gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "qr_tmerge");
      for (size_t q = 0; q < jobs.size(); ++q) {
        const Aux& a = aux[q];
        const int nb = blk_cols(a, G * s + h);
        if (nb <= 0) continue;
        gb.add(h * NBO, nb, h * NBO, -1.0, ts(a, s), 1, NS, a.Y, 1, NS, 0.0, ts(a, s) + int64_t(h) * NBO * NS, 1, NS);
      }
      gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "qr_tmerge");
    }
    // the group applied to the trailing columns
    apply(s, true, false, "qr_trail", grp_cols, [&](const QRJob& J, const Aux& a, int ss, int& j0, int& j1) {
      j0 = NS * ss + grp_cols(a, ss);
      j1 = J.n;
    });
  }
  // ---- thin Q by backward accumulation over the groups: Q2 -= V_s (T_s (V_s^T Q2))
  std::vector<Fill2D> fills;
  for (size_t q = 0; q < jobs.size(); ++q)
    if (jobs[q].Q && aux[q].kq > 0) fills.push_back(Fill2D{jobs[q].Q, jobs[q].ldq, jobs[q].m, aux[q].kq, 1.0});
  if (fills.empty()) return;
  fill2d_batched(ctx, fills);
  for (int s = max_grp - 1; s >= 0; --s)
    apply(s, false, true, "qr_formq", grp_cols, [&](const QRJob&, const Aux& a, int ss, int& j0, int& j1) {
      j0 = NS * ss;
      j1 = a.kq;
    });
}

}  // namespace

void qr_batched(Ctx& ctx, std::vector<QRJob>& jobs) {
  WsScope scope(ctx);  // tau, T factors and the blocks' V are needed only here
  {
    int mm = 0, nn = 0;
    size_t bytes = 0;
    for (const auto& J : jobs) {
      mm = std::max(mm, J.m);
      nn = std::max(nn, J.n);
      bytes += size_t(std::max(J.m, 0)) * size_t(std::max(J.n, 0)) * sizeof(double);
    }
    // One workgroup per matrix pays off only when the batch occupies the chip (or a
    // matrix is a single panel): a few matrices run their column steps on a few CUs,
    // while the blocked path spreads every update over all of them.
    const bool fits = mm <= QR_SMALL_MAX_M && bytes <= ctx.qr_small_bytes;
    const bool wide = jobs.size() >= QR_SMALL_MIN_BATCH || nn <= 16;
    if ((ctx.qr_path == 0 && fits && wide) || (ctx.qr_path == 3 && mm <= QR_SMALL_MAX_M)) {
      qr_small_batched(ctx, jobs);
      return;
    }
  }
  {
    // register-resident blocks (panels of 16) up to 6*512 rows, taken in groups; else
    // 64-column blocks of streaming panels (width 32)
    int max_m = 0;
    for (const auto& J : jobs) max_m = std::max(max_m, J.m);
    if (max_m <= 6 * RT && ctx.qr_path != 2) {
      qr_groups(ctx, jobs, ctx.qr_group);
      return;
    }
  }
  struct Aux {
    double* tau;
    double* T;   // inner panels: npanel * NBP * NBP
    double* TO;  // blocks: nblock * NBO * NBO
    double* G;   // NBO x NBO
    double* W;   // NBO x n
    double* W2;  // NBO x n: T W (or T^T W)
    double* VX;  // the blocks' V with explicit unit diagonal and zeros above, block b
                 // ((m - 64 b) x 64, ld m - 64 b) at vx(b); one buffer for all blocks
                 // when no Q is formed, else every block's kept for the formation of Q
    bool vx_all;
    int kq, nblock;
  };
  std::vector<Aux> aux(jobs.size());
  int max_blocks = 0;
  const int NB = NBP;
  for (size_t q = 0; q < jobs.size(); ++q) {
    const QRJob& J = jobs[q];
    TTR_REQUIRE(J.m >= 0 && J.n >= 0 && J.lda >= std::max(1, J.m), ST_INVALID_ARG, "qr: bad shape");
    Aux& a = aux[q];
    a.kq = std::min(J.m, J.n);
    a.nblock = (a.kq + NBO - 1) / NBO;
    max_blocks = std::max(max_blocks, a.nblock);
    const int npanel = (a.kq + NB - 1) / NB;
    a.tau = ctx.ws.alloc_n<double>(std::max(1, J.n));
    a.T = ctx.ws.alloc_n<double>(size_t(std::max(1, npanel)) * NBP * NBP);
    a.TO = ctx.ws.alloc_n<double>(size_t(std::max(1, a.nblock)) * NBO * NBO);
    a.G = ctx.ws.alloc_n<double>(size_t(NBO) * NBO);
    a.W = ctx.ws.alloc_n<double>(size_t(NBO) * std::max(1, J.n));
    a.W2 = ctx.ws.alloc_n<double>(size_t(NBO) * std::max(1, J.n));
    a.vx_all = J.Q != nullptr;
    a.VX = ctx.ws.alloc_n<double>(
        a.vx_all ? size_t(NBO) * (size_t(a.nblock) * J.m - size_t(NBO / 2) * a.nblock * std::max(0, a.nblock - 1))
                 : size_t(std::max(1, J.m)) * NBO);
  }
  GemmBatch gb;
  auto block_cols = [&](const Aux& a, int b) { return std::min(NBO, a.kq - b * NBO); };
  auto vx = [&](const QRJob& J, const Aux& a, int b) {
    return a.VX + (a.vx_all ? int64_t(NBO) * (int64_t(b) * J.m - int64_t(NBO / 2) * b * (b - 1)) : 0);
  };
  // explicit V of block b (rows b*NBO.., ld m2) for every job that has it
  auto copy_vx = [&](int b) {
    std::vector<Copy2D> cp;
    for (size_t q = 0; q < jobs.size(); ++q) {
      const QRJob& J = jobs[q];
      const Aux& a = aux[q];
      if (b >= a.nblock) continue;
      const int c0 = b * NBO, m2 = J.m - c0;
      cp.push_back(copy_job(J.A + c0 + J.lda * c0, 1, J.lda, vx(J, a, b), 1, m2, m2, block_cols(a, b), 1.0, 3));
    }
    copy2d_batched(ctx, cp);
  };
  // ---- factorisation
  for (int b = 0; b < max_blocks; ++b) {
    for (int ip = 0; ip < NBO / NB; ++ip) {
      std::vector<PanelJob> pj;
      for (size_t q = 0; q < jobs.size(); ++q) {
        const Aux& a = aux[q];
        if (b >= a.nblock) continue;
        const int c = b * NBO + ip * NB, cend = b * NBO + block_cols(a, b);
        if (c >= cend) continue;
        pj.push_back(PanelJob{jobs[q].A, jobs[q].lda, jobs[q].m, c, std::min(NB, cend - c), a.tau,
                              a.T + size_t(c / NB) * NBP * NBP});
      }
      if (pj.empty()) continue;
      const PanelJob* d = ctx.upload(pj);
      double pfl = 0.0, pby = 0.0;
      for (const auto& j : pj) {
        pfl += 4.0 * double(j.m - j.c0) * j.nb * j.nb;  // Householder panel (2 m nb^2) + T
        pby += 16.0 * double(j.m - j.c0) * j.nb;
      }
      {
        ProfScope ps(ctx, "qr_panel", pfl, pby);
        launch_panel(ctx, d, int(pj.size()));
      }
      // apply the panel to the rest of its block: A2 -= V (T^T (V^T A2)); the
      // triangular factor multiplies the short side (nb x n2), not V (m2 x nb)
      for (int stage = 0; stage < 3; ++stage) {
        for (size_t q = 0; q < jobs.size(); ++q) {
          const QRJob& J = jobs[q];
          const Aux& a = aux[q];
          if (b >= a.nblock) continue;
          const int c = b * NBO + ip * NB, cend = b * NBO + block_cols(a, b);
          if (c >= cend) continue;
          const int nb = std::min(NB, cend - c);
          const int m2 = J.m - c, n2 = cend - c - nb;
          if (n2 <= 0) continue;
          double* V = J.A + c + J.lda * c;
          double* A2 = J.A + c + J.lda * (c + nb);
          const double* T = a.T + size_t(c / NB) * NBP * NBP;
          if (stage == 0)  // W = V^T A2  (nb x n2)
            gb.add(nb, n2, m2, 1.0, V, J.lda, 1, A2, 1, J.lda, 0.0, a.W, 1, NBO, AM_VT_UNIT);
          else if (stage == 1)  // W2 = T^T W  (nb x n2)
            gb.add(nb, n2, nb, 1.0, T, NBP, 1, a.W, 1, NBO, 0.0, a.W2, 1, NBO);
          else  // A2 -= V W2
            gb.add(m2, n2, nb, -1.0, V, 1, J.lda, a.W2, 1, NBO, 1.0, A2, 1, J.lda, AM_V_UNIT);
        }
        gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "qr_panel_upd");
      }
    }
    // merge the block's reflectors: G = V^T V, T from (G, tau)
    copy_vx(b);
    std::vector<TMergeJob> tb;
    for (size_t q = 0; q < jobs.size(); ++q) {
      const QRJob& J = jobs[q];
      const Aux& a = aux[q];
      if (b >= a.nblock) continue;
      const int c0 = b * NBO, m2 = J.m - c0, nbo = block_cols(a, b);
      gb.add(nbo, nbo, m2, 1.0, vx(J, a, b), m2, 1, vx(J, a, b), 1, m2, 0.0, a.G, 1, NBO);
      tb.push_back(TMergeJob{a.G, a.T + size_t(c0 / NB) * NBP * NBP, a.TO + size_t(b) * NBO * NBO, nbo, NB, NBO, NBO, 0});
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "qr_vtv");
    launch_merge_t(ctx, tb);
    // trailing update A2 -= V (T^T (V^T A2))
    for (int stage = 0; stage < 3; ++stage) {
      for (size_t q = 0; q < jobs.size(); ++q) {
        const QRJob& J = jobs[q];
        const Aux& a = aux[q];
        if (b >= a.nblock) continue;
        const int c0 = b * NBO, nbo = block_cols(a, b);
        const int m2 = J.m - c0, n2 = J.n - c0 - nbo;
        if (n2 <= 0) continue;
        double* A2 = J.A + c0 + J.lda * (c0 + nbo);
        const double* T = a.TO + size_t(b) * NBO * NBO;
        if (stage == 0)  // W = V^T A2  (nbo x n2)
          gb.add(nbo, n2, m2, 1.0, vx(J, a, b), m2, 1, A2, 1, J.lda, 0.0, a.W, 1, NBO);
        else if (stage == 1)  // W2 = T^T W  (nbo x n2)
          gb.add(nbo, n2, nbo, 1.0, T, NBO, 1, a.W, 1, NBO, 0.0, a.W2, 1, NBO);
        else  // A2 -= V W2
          gb.add(m2, n2, nbo, -1.0, vx(J, a, b), 1, m2, a.W2, 1, NBO, 1.0, A2, 1, J.lda);
      }
      gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "qr_trail");
    }
  }
  // ---- thin Q by backward accumulation over the blocks: Q2 -= V (T (V^T Q2))
  std::vector<Fill2D> fills;
  for (size_t q = 0; q < jobs.size(); ++q)
    if (jobs[q].Q && aux[q].kq > 0) fills.push_back(Fill2D{jobs[q].Q, jobs[q].ldq, jobs[q].m, aux[q].kq, 1.0});
  if (fills.empty()) return;
  fill2d_batched(ctx, fills);
  // (every job here forms Q, so its blocks' V are all kept from the factorisation)
  for (int b = max_blocks - 1; b >= 0; --b) {
    for (int stage = 0; stage < 3; ++stage) {
      for (size_t q = 0; q < jobs.size(); ++q) {
        const QRJob& J = jobs[q];
        const Aux& a = aux[q];
        if (!J.Q || b >= a.nblock) continue;
        const int c0 = b * NBO, nbo = block_cols(a, b);
        const int m2 = J.m - c0, n2 = a.kq - c0;
        double* Q2 = J.Q + c0 + J.ldq * c0;
        const double* T = a.TO + size_t(b) * NBO * NBO;
        if (stage == 0)  // W = V^T Q2  (nbo x n2)
          gb.add(nbo, n2, m2, 1.0, vx(J, a, b), m2, 1, Q2, 1, J.ldq, 0.0, a.W, 1, NBO);
        else if (stage == 1)  // W2 = T W  (nbo x n2)
          gb.add(nbo, n2, nbo, 1.0, T, 1, NBO, a.W, 1, NBO, 0.0, a.W2, 1, NBO);
        else  // Q2 -= V W2
          gb.add(m2, n2, nbo, -1.0, vx(J, a, b), 1, m2, a.W2, 1, NBO, 1.0, Q2, 1, J.ldq);
      }
      gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "qr_formq");
    }
  }
}

}  // namespace ttr
