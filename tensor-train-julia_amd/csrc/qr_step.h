// One Householder column step of a 16-column panel held in registers, shared by
// the one-launch QR (qr_small.hip) and the panel kernel of the blocked QR
// (qr.hip).  LAPACK conventions (dlarfg/dlarft):
//   beta = -sign(alpha) ||(alpha, x)||, tau = (beta - alpha)/beta,
//   v = (1, x/(alpha - beta)), tau = 0 when x == 0.
#pragma once
#include <hip/hip_runtime.h>

namespace ttr {
namespace qrstep {

constexpr int SNB = 16;        // panel width
constexpr int LDV = SNB + 1;   // LDS row stride of 16-wide arrays (odd: conflict-free column reads)

template <int NT>
struct PanelShared16 {
  double Ts[SNB * LDV];         // T of the current panel, Ts[r * LDV + c] = T(r, c)
  double red[2][NT / 64][SNB];  // per-wave column sums (double-buffered by step parity)
  double rowj[2][SNB];          // the pivot row of the panel before the step
};

// Cross-lane moves of a double on the VALU (no LDS crossbar round trip): DPP row and
// quad permutations, and gfx950's lane-half swaps between two registers.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
constexpr int DPP_XOR1 = 0xB1;         // quad_perm [1, 0, 3, 2]
constexpr int DPP_XOR2 = 0x4E;         // quad_perm [2, 3, 0, 1]
constexpr int DPP_XOR3 = 0x1B;         // quad_perm [3, 2, 1, 0]
constexpr int DPP_HALF_MIRROR = 0x141; // lane l of each 8 <- lane 7 - l, i.e. l ^ 7
constexpr int DPP_ROR8 = 0x128;        // row_ror:8, i.e. l ^ 8 within rows of 16

__device__ __forceinline__ double readlane_d(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

// Lanes 0-31: x_l + x_{l+32}; lanes 32-63: y_{l-32} + y_l (v_permlane32_swap).
__device__ __forceinline__ double swap32_add(double x, double y) {
  const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(x), __double2loint(y), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(x), __double2hiint(y), false, false);
  return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}
// Rows 0, 2 (of 16 lanes): x_l + x_{l+16}; rows 1, 3: y_{l-16} + y_l (v_permlane16_swap).
__device__ __forceinline__ double swap16_add(double x, double y) {
  const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(x), __double2loint(y), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(x), __double2hiint(y), false, false);
  return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}

// The 16 panel sums of a wave: lane l ends with the wave total of column
// (l >> 2) & 15 (lanes 4c .. 4c+3 all hold column c).  A reduce-scatter: at each
// level a lane keeps the half of its columns its partner lane does not, and adds the
// partner's values of that half (partners at lane distance 32, 16, 8, 4; then 2, 1
// for the plain sum); the moves are the VALU ones above.
__device__ __forceinline__ double butterfly16(const double (&acc)[SNB], int lane) {
  double a8[8], a4[4], a2[2];
  const bool h3 = lane & 8, h2 = lane & 4;
#pragma unroll
  for (int c = 0; c < 8; ++c) a8[c] = swap32_add(acc[c], acc[c + 8]);
#pragma unroll
  for (int c = 0; c < 4; ++c) a4[c] = swap16_add(a8[c], a8[c + 4]);
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const double send = h3 ? a4[c] : a4[c + 2];
    const double keep = h3 ? a4[c + 2] : a4[c];
    a2[c] = keep + dpp_d<DPP_ROR8>(send);
  }
  double a1;
  {
    const double send = h2 ? a2[0] : a2[1];
    const double keep = h2 ? a2[1] : a2[0];
    a1 = keep + dpp_d<DPP_XOR3>(dpp_d<DPP_HALF_MIRROR>(send));  // lane l ^ 4
  }
  a1 += dpp_d<DPP_XOR2>(a1);
  a1 += dpp_d<DPP_XOR1>(a1);
  return a1;
}

// Householder step JJ of a 16-column panel held in registers: threads tid < NT
// hold panel rows tid + NT q (R of them); threads past NT hold no rows and only
// meet the barrier.  One workgroup barrier per column: the 16 sums
// s_c = x^T P(:, c) of the column below the pivot (x) against every panel column
// are reduced within each wave (butterfly reduce-scatter) and written to LDS;
// after the barrier every wave adds the waves' partial sums itself (two values per
// lane and two lane swaps, the same operations in every wave, so every wave has
// the same bits), forms the reflector's scalars (the squared norm below the pivot
// is s_JJ) and d_c = v^T P(:, c) over rows >= JJ (v_JJ = 1) in lane c, and reads
// the 16 values into scalar registers: for c > JJ they apply the reflector to the
// remaining columns, for c < JJ they are the entries V^T v of T.  The steps are
// instantiated per column so that every register of the panel is addressed
// statically (a run-time column loop with selects measured slower).
template <int R, int NT, int JJ>
struct PanelStep16 {
  __device__ static __forceinline__ void run(double (&P)[R][SNB], PanelShared16<NT>& sh, double* tau_out, int nb,
                                             int m2, int tid, int lane, int wave) {
    constexpr int NW = NT / 64;
    static_assert(NW == 4 || NW == 8, "four or eight waves hold the panel");
    if (JJ < nb) {
      constexpr int buf = JJ & 1;
      const bool pt = tid < NT;
      // s_c = sum_{i > JJ} x_i P[i][c] (s_JJ = sigma, the squared norm below the pivot)
      double acc[SNB];
#pragma unroll
      for (int c = 0; c < SNB; ++c) acc[c] = 0.0;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = tid + NT * q;
        if (pt && i > JJ && i < m2) {
          const double x = P[q][JJ];
#pragma unroll
          for (int c = 0; c < SNB; ++c) acc[c] += x * P[q][c];
        }
      }
      if (pt) {
        const double ws = butterfly16(acc, lane);
        if ((lane & 3) == 0) sh.red[buf][wave][lane >> 2] = ws;
      }
      if (tid == JJ) {
#pragma unroll
        for (int c = 0; c < SNB; ++c) sh.rowj[buf][c] = P[0][c];
      }
      __syncthreads();
      // lane l holds column c = l % 16 of waves l / 16 (and l / 16 + 4), summed over
      // the four lane groups: every lane ends with the total of its column
      const int cl = lane & 15, g = lane >> 4;
      double v = sh.red[buf][g][cl];
      if (NW == 8) v += sh.red[buf][g + 4][cl];
      v = swap32_add(v, v);
      v = swap16_add(v, v);
      const double sg = readlane_d(v, JJ);
      const double al = sh.rowj[buf][JJ];
      double tau = 0.0, beta = al, scal = 1.0;
      if (sg != 0.0) {
        beta = -copysign(sqrt(al * al + sg), al);
        tau = (beta - al) / beta;
        scal = 1.0 / (al - beta);
      }
      const double dl = sh.rowj[buf][cl] + scal * v;
      double dd[SNB];
#pragma unroll
      for (int c = 0; c < SNB; ++c) dd[c] = readlane_d(dl, c);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = tid + NT * q;
        if (pt && i > JJ && i < m2) {
          const double x = P[q][JJ] * scal;
          P[q][JJ] = x;
#pragma unroll
          for (int c = JJ + 1; c < SNB; ++c) P[q][c] -= tau * x * dd[c];
        }
      }
      if (tid == JJ) {
        P[0][JJ] = beta;
#pragma unroll
        for (int c = JJ + 1; c < SNB; ++c) P[0][c] -= tau * dd[c];
      }
      if (tid == 0) tau_out[JJ] = tau;
      // T(0:JJ, JJ) = -tau T(0:JJ, 0:JJ) d(0:JJ),  T(JJ, JJ) = tau
      if (tid < JJ) {
        double t = 0.0;
#pragma unroll
        for (int qq = 0; qq < JJ; ++qq)
          if (qq >= tid) t += sh.Ts[tid * LDV + qq] * dd[qq];
        sh.Ts[tid * LDV + JJ] = -tau * t;
      }
      if (tid == JJ) sh.Ts[JJ * LDV + JJ] = tau;
    }
    PanelStep16<R, NT, JJ + 1>::run(P, sh, tau_out, nb, m2, tid, lane, wave);
  }
};

template <int R, int NT>
struct PanelStep16<R, NT, SNB> {
  __device__ static __forceinline__ void run(double (&)[R][SNB], PanelShared16<NT>&, double*, int, int, int, int,
                                             int) {}
};

}  // namespace qrstep
}  // namespace ttr
