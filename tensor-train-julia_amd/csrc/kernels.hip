// Data-movement, RNG and small elementwise kernels of libttround.
#include "ttr_internal.h"

namespace ttr {

namespace {

__device__ __forceinline__ float bf16_to_f32_(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16_(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

__device__ __forceinline__ void store_prec(void* p, int64_t i, double v, int prec) {
  if (prec == 0) static_cast<double*>(p)[i] = v;
  else if (prec == 1) static_cast<float*>(p)[i] = float(v);
  else static_cast<uint16_t*>(p)[i] = f32_to_bf16_(float(v));
}

template <class T>
__device__ __forceinline__ double ld_as_d(const T* p, int64_t i);
template <> __device__ __forceinline__ double ld_as_d<double>(const double* p, int64_t i) { return p[i]; }
template <> __device__ __forceinline__ double ld_as_d<float>(const float* p, int64_t i) { return p[i]; }
template <> __device__ __forceinline__ double ld_as_d<uint16_t>(const uint16_t* p, int64_t i) { return bf16_to_f32_(p[i]); }

template <class T>
__device__ __forceinline__ void st_from_d(T* p, int64_t i, double v);
template <> __device__ __forceinline__ void st_from_d<double>(double* p, int64_t i, double v) { p[i] = v; }
template <> __device__ __forceinline__ void st_from_d<float>(float* p, int64_t i, double v) { p[i] = float(v); }
template <> __device__ __forceinline__ void st_from_d<uint16_t>(uint16_t* p, int64_t i, double v) { p[i] = f32_to_bf16_(float(v)); }

__device__ __forceinline__ int find_job(const int32_t* prefix, int njobs, int bid) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= bid) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter-based.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
    const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Standard normal draw number e of core k: counter (e_lo, e_hi, k, 0x5EED),
// key = seed; two 53-bit uniforms, Box-Muller cosine branch, in FP64.
__device__ __forceinline__ double rng_normal(uint64_t seed, int core, uint64_t e, uint32_t tag = 0x5EEDu) {
  const uint4 x = philox4x32_10(make_uint4(uint32_t(e), uint32_t(e >> 32), uint32_t(core), tag),
                                uint32_t(seed), uint32_t(seed >> 32));
  const uint64_t a = ((uint64_t(x.x) << 32) | x.y) >> 11;
  const uint64_t b = ((uint64_t(x.z) << 32) | x.w) >> 11;
  const double u1 = (double(a) + 0.5) * 0x1.0p-53;
  const double u2 = double(b) * 0x1.0p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// ---------------------------------------------------------------- kernels

struct Copy2DDev {
  Copy2D j;
  int32_t tile0, tiles_n, tiles_m;
};

constexpr int SMALL_ARG_JOBS = 8;  // jobs of a copy, fill, gather or scaling launch passed in its arguments

template <bool INL>
__global__ __launch_bounds__(256) void copy2d_kernel(const Copy2DDev* __restrict__ jobs, const int32_t* __restrict__ prefix,
                                                     int njobs, ArgJobs<Copy2DDev, SMALL_ARG_JOBS, INL> aj) {
  __shared__ double tile[32][33];
  int jid = 0;
  if constexpr (INL) {
    jobs = aj.v;
    for (int q = 1; q < njobs; ++q)
      if (jobs[q].tile0 <= int(blockIdx.x)) jid = q;
  } else {
    jid = find_job(prefix, njobs, blockIdx.x);
  }
  const Copy2DDev D = jobs[jid];
  const Copy2D& J = D.j;
  int t = blockIdx.x - D.tile0;
  const int per_b = D.tiles_m * D.tiles_n;
  const int b = t / per_b;
  t -= b * per_b;
  const int i0 = (t / D.tiles_n) * 32, j0 = (t % D.tiles_n) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const double* in = J.in + b * J.ibs;
  double* out = J.out + b * J.obs;
  const bool in_i_fast = (J.irs == 1);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int ii = in_i_fast ? tx : ty + 8 * q;
    const int jj = in_i_fast ? ty + 8 * q : tx;
    const int gi = i0 + ii, gj = j0 + jj;
    if (gi < J.M && gj < J.N) {
      const bool zero = (J.tri == 1 && gi > gj) || ((J.tri == 2 || J.tri == 3) && gi < gj);
      tile[ii][jj] = zero ? 0.0 : (J.tri == 3 && gi == gj) ? 1.0 : J.alpha * in[gi * J.irs + gj * J.ics];
    }
  }
  __syncthreads();
  const bool out_i_fast = (J.ors == 1);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int ii = out_i_fast ? tx : ty + 8 * q;
    const int jj = out_i_fast ? ty + 8 * q : tx;
    const int gi = i0 + ii, gj = j0 + jj;
    if (gi < J.M && gj < J.N) out[gi * J.ors + gj * J.ocs] = tile[ii][jj];
  }
}

template <bool INL>
__global__ void fill2d_kernel(const Fill2D* __restrict__ jobs, ArgJobs<Fill2D, SMALL_ARG_JOBS, INL> aj) {
  if constexpr (INL) jobs = aj.v;
  const Fill2D J = jobs[blockIdx.y];
  const int64_t tot = int64_t(J.M) * J.N;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < tot; e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = e % J.M, c = e / J.M;
    J.out[i + c * J.ld] = (i == c) ? J.diag : 0.0;
  }
}

template <bool INL>
__global__ void gather_cols_kernel(const GatherCols* __restrict__ jobs, ArgJobs<GatherCols, SMALL_ARG_JOBS, INL> aj) {
  if constexpr (INL) jobs = aj.v;
  const GatherCols J = jobs[blockIdx.y];
  const int64_t tot = int64_t(J.M) * J.ncol;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < tot; e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = e % J.M, c = e / J.M;
    J.out[i + c * J.ldo] = J.in[i + int64_t(J.idx[c]) * J.ldi];
  }
}

template <class T, bool INL>
__global__ __launch_bounds__(256) void pow2_kernel(const Pow2Job* __restrict__ jobs,
                                                   ArgJobs<Pow2Job, SMALL_ARG_JOBS, INL> aj) {
  if constexpr (INL) jobs = aj.v;
  const Pow2Job J = jobs[blockIdx.x];
  T* X = static_cast<T*>(J.X);
  // one column per wave at a time (no 64-bit division per element)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double m = 0.0;
  for (int c = wave; c < J.N; c += 4)
    for (int i = lane; i < J.M; i += 64) m = fmax(m, fabs(ld_as_d<T>(X, i + int64_t(c) * J.ld)));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  int e2 = 0;
  if (m > 0.0 && isfinite(m)) e2 = ilogb(m);
  const double s = ldexp(1.0, -e2);
  for (int c = wave; c < J.N; c += 4)
    for (int i = lane; i < J.M; i += 64) {
      const int64_t o = i + int64_t(c) * J.ld;
      st_from_d<T>(X, o, ld_as_d<T>(X, o) * s);
    }
  if (threadIdx.x == 0) J.exp_out[0] = (J.exp_in ? J.exp_in[0] : 0) + e2;
}

__global__ void sketch_gen_kernel(const SketchGen* __restrict__ jobs, uint64_t seed, int prec) {
  const SketchGen J = jobs[blockIdx.y];
  const int64_t tot = int64_t(J.n) * J.bvr * J.bvl * J.p;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < tot; e += int64_t(gridDim.x) * blockDim.x) {
    // e is the reference (randn!) order index: beta + bvr*(z + n*(b + bvl*j))
    const double g = rng_normal(seed, J.core, uint64_t(e));
    const int64_t beta = e % J.bvr;
    const int64_t rest = e / J.bvr;
    const int64_t z = rest % J.n;
    const int64_t col = rest / J.n;  // b + bvl*j
    if (J.out) store_prec(J.out, z + J.n * (beta + int64_t(J.bvr) * col), J.scale * g, prec);
    if (J.ref_out) J.ref_out[e] = J.ref_scale * g;
  }
}

__global__ void randn_core_kernel(double* out, int n, int r, int r2, int core, uint64_t seed) {
  const int64_t tot = int64_t(n) * r * r2;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < tot; e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = e % n;
    const int64_t rest = e / n;
    const int64_t a = rest % r;
    const int64_t b = rest / r;
    out[a + int64_t(r) * (i + int64_t(n) * b)] = rng_normal(seed, core, uint64_t(e), 0x7A11u);
  }
}

__global__ void sketch_draws_kernel(double* out, int64_t count, int core, uint64_t seed) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < count; e += int64_t(gridDim.x) * blockDim.x)
    out[e] = rng_normal(seed, core, uint64_t(e));
}

__global__ void sketch_pack_kernel(const double* __restrict__ ref, void* out, int n, int bvr, int bvl, int p,
                                   double scale, int prec) {
  const int64_t tot = int64_t(n) * bvr * bvl * p;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < tot; e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t beta = e % bvr;
    const int64_t rest = e / bvr;
    const int64_t z = rest % n;
    const int64_t col = rest / n;
    store_prec(out, z + n * (beta + int64_t(bvr) * col), scale * ref[e], prec);
  }
}

__global__ void zero_kernel(uint32_t* p, int64_t words) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < words; e += int64_t(gridDim.x) * blockDim.x)
    p[e] = 0u;
}

struct SumJob {
  const double* const* ins;
  int nin;
};

__global__ void sum_kernel(double* out, const double* const* __restrict__ ins, int nin, int64_t count) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < count; e += int64_t(gridDim.x) * blockDim.x) {
    double s = 0.0;
    for (int t = 0; t < nin; ++t) s += ins[t][e];
    out[e] = s;
  }
}

// one thread per output entry (consecutive threads: consecutive left ranks, so
// the stores are contiguous); 32-bit index arithmetic (the host bounds every
// core below 2^31 entries)
__global__ void kron_core_kernel(const KronCore* __restrict__ jobs) {
  const KronCore J = jobs[blockIdx.y];
  const uint32_t RA = uint32_t(J.RA), n = uint32_t(J.n);
  const uint32_t tot = RA * n * uint32_t(J.RB);
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += gridDim.x * blockDim.x) {
    uint32_t a = e % RA;
    const uint32_t rest = e / RA;
    const uint32_t i = rest % n;
    uint32_t b = rest / n;
    double v = 1.0;
    for (int f = 0; f < J.m; ++f) {
      const uint32_t ra = uint32_t(J.ra[f]), rb = uint32_t(J.rb[f]);
      const uint32_t af = a % ra, bf = b % rb;
      a /= ra;
      b /= rb;
      v *= J.in[f][af + ra * (i + n * bf)];
    }
    J.out[e] = v;
  }
}

__global__ void block_put_kernel(const BlockPut* __restrict__ jobs) {
  const BlockPut J = jobs[blockIdx.y];
  const int64_t tot = int64_t(J.ra) * J.n * J.rb;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < tot; e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t a = e % J.ra;
    const int64_t rest = e / J.ra;
    const int64_t i = rest % J.n;
    const int64_t b = rest / J.n;
    J.out[(J.a0 + a) + int64_t(J.RA) * (i + int64_t(J.n) * (J.b0 + b))] = J.scale * J.in[e];
  }
}

__global__ void mpo_kernel(const MpoCore* __restrict__ jobs) {
  const MpoCore J = jobs[blockIdx.y];
  const int64_t R1 = int64_t(J.Ra) * J.rx, R2 = int64_t(J.Ra2) * J.rx2;
  const int64_t tot = R1 * J.n * R2;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < tot; e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r1 = e % R1;
    const int64_t rest = e / R1;
    const int64_t i = rest % J.n;
    const int64_t r2 = rest / J.n;
    const int64_t al = r1 % J.Ra, nu = r1 / J.Ra;
    const int64_t al2 = r2 % J.Ra2, nu2 = r2 / J.Ra2;
    double s = 0.0;
    for (int j = 0; j < J.n; ++j)
      s += J.A[i + J.n * (j + int64_t(J.n) * (al + int64_t(J.Ra) * al2))] * J.x[nu + int64_t(J.rx) * (j + int64_t(J.n) * nu2)];
    J.y[e] = s;
  }
}

// Local action of an operator core on (mode, bond) pairs, one thread per fiber:
// out[(om, ob)] = scale * sum_{im, ib} H(om, ob, im, ib) in[(im, ib)].  The
// (n*rout) x (n*rin) matrix is staged in LDS with row pitch MAXO, zero padded to
// MAXO outputs and to a multiple of 8 inputs, so the accumulation loop carries
// no guards and reads the matrix as aligned vectors; a thread keeps its MAXO
// outputs in registers and issues its inputs 8 loads at a time.
template <int MAXO, class TI, class TO>
__global__ __launch_bounds__(256) void op_apply_kernel(const OpApply* __restrict__ jobs) {
  const OpApply J = jobs[blockIdx.y];
  __shared__ __attribute__((aligned(16))) double M[64 * MAXO];
  __shared__ int64_t offi[64], offo[64];
  const int n = J.n, din = J.n * J.rin, dout = J.n * J.rout;
  const int din8 = (din + 7) & ~7;
  for (int e = threadIdx.x; e < din8 * MAXO; e += blockDim.x) {
    const int o = e % MAXO, i = e / MAXO;
    M[e] = (o < dout && i < din)
               ? J.H[(o % n) * J.h_om + (o / n) * J.h_ob + (i % n) * J.h_im + (i / n) * J.h_ib]
               : 0.0;
  }
  for (int e = threadIdx.x; e < 64; e += blockDim.x) {
    offi[e] = e < din ? (e % n) * J.is_m + (e / n) * J.is_b : 0;
    offo[e] = e < dout ? (e % n) * J.os_m + (e / n) * J.os_b : 0;
  }
  __syncthreads();
  const int64_t nf = int64_t(J.X) * J.Y;
  for (int64_t f = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; f < nf; f += int64_t(gridDim.x) * blockDim.x) {
    const int64_t x = f % J.X, y = f / J.X;
    const TI* in = static_cast<const TI*>(J.in) + x * J.is_x + y * J.is_y;
    double acc[MAXO];
#pragma unroll
    for (int o = 0; o < MAXO; ++o) acc[o] = 0.0;
    for (int i0 = 0; i0 < din8; i0 += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = (i0 + u < din) ? ld_as_d<TI>(in, offi[i0 + u]) : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double* m = M + (i0 + u) * MAXO;
#pragma unroll
        for (int o = 0; o < MAXO; ++o) acc[o] += m[o] * v[u];
      }
    }
    TO* out = static_cast<TO*>(J.out) + x * J.os_x + y * J.os_y;
#pragma unroll
    for (int o = 0; o < MAXO; ++o)
      if (o < dout) st_from_d<TO>(out, offo[o], acc[o] * J.scale);
  }
}

}  // namespace

// ---------------------------------------------------------------- host launchers

void copy2d_batched(Ctx& ctx, const std::vector<Copy2D>& jobs) {
  std::vector<Copy2DDev> dj;
  std::vector<int32_t> prefix;
  int64_t tiles = 0;
  for (const auto& j : jobs) {
    if (j.M <= 0 || j.N <= 0 || j.nb <= 0) continue;
    Copy2DDev d{};
    d.j = j;
    d.tiles_m = (j.M + 31) / 32;
    d.tiles_n = (j.N + 31) / 32;
    d.tile0 = int32_t(tiles);
    prefix.push_back(int32_t(tiles));
    tiles += int64_t(d.tiles_m) * d.tiles_n * j.nb;
    dj.push_back(d);
  }
  if (dj.empty()) return;
  TTR_REQUIRE(tiles < (int64_t(1) << 31), ST_INVALID_ARG, "copy too large");
  const dim3 g{unsigned(tiles)}, b{256u};
  if (arg_jobs_fit<SMALL_ARG_JOBS>(ctx, dj)) {
    TTR_LAUNCH(ctx, copy2d_kernel<true>, g, b, 0, nullptr, nullptr, int(dj.size()), arg_jobs<SMALL_ARG_JOBS>(dj));
  } else {
    const Copy2DDev* d = ctx.upload(dj);
    const int32_t* pre = ctx.upload(prefix);
    TTR_LAUNCH(ctx, copy2d_kernel<false>, g, b, 0, d, pre, int(dj.size()), (ArgJobs<Copy2DDev, SMALL_ARG_JOBS, false>{}));
  }
  TTR_HIP(hipGetLastError());
}

// one workgroup row per job (blockIdx.y), up to 1024 workgroups along x
template <class J>
static void launch_y(Ctx& ctx, void (*kern)(const J*), const std::vector<J>& jobs, int64_t max_elems) {
  int64_t bx = (max_elems + 255) / 256;
  if (bx > 1024) bx = 1024;
  if (bx < 1) bx = 1;
  for (size_t s = 0; s < jobs.size(); s += 65535) {
    const size_t cnt = std::min<size_t>(65535, jobs.size() - s);
    const J* d = ctx.upload(jobs.data() + s, cnt);
    TTR_LAUNCH(ctx, kern, dim3(unsigned(bx), unsigned(cnt)), dim3(256), 0, d);
    TTR_HIP(hipGetLastError());
  }
}

// the same for a kernel that also takes up to SMALL_ARG_JOBS jobs in its arguments
template <class J>
static void launch_y(Ctx& ctx, void (*kern)(const J*, ArgJobs<J, SMALL_ARG_JOBS, false>),
                     void (*kern_args)(const J*, ArgJobs<J, SMALL_ARG_JOBS, true>), const std::vector<J>& jobs,
                     int64_t max_elems) {
  if (jobs.empty()) return;
  int64_t bx = (max_elems + 255) / 256;
  if (bx > 1024) bx = 1024;
  if (bx < 1) bx = 1;
  if (arg_jobs_fit<SMALL_ARG_JOBS>(ctx, jobs)) {
    TTR_LAUNCH(ctx, kern_args, dim3(unsigned(bx), unsigned(jobs.size())), dim3(256), 0, nullptr,
               arg_jobs<SMALL_ARG_JOBS>(jobs));
    TTR_HIP(hipGetLastError());
    return;
  }
  for (size_t s = 0; s < jobs.size(); s += 65535) {
    const size_t cnt = std::min<size_t>(65535, jobs.size() - s);
    const J* d = ctx.upload(jobs.data() + s, cnt);
    TTR_LAUNCH(ctx, kern, dim3(unsigned(bx), unsigned(cnt)), dim3(256), 0, d, (ArgJobs<J, SMALL_ARG_JOBS, false>{}));
    TTR_HIP(hipGetLastError());
  }
}

void fill2d_batched(Ctx& ctx, const std::vector<Fill2D>& jobs) {
  int64_t mx = 0;
  for (auto& j : jobs) mx = std::max<int64_t>(mx, int64_t(j.M) * j.N);
  launch_y(ctx, fill2d_kernel<false>, fill2d_kernel<true>, jobs, mx);
}

void gather_cols_batched(Ctx& ctx, const std::vector<GatherCols>& jobs) {
  int64_t mx = 0;
  for (auto& j : jobs) mx = std::max<int64_t>(mx, int64_t(j.M) * j.ncol);
  launch_y(ctx, gather_cols_kernel<false>, gather_cols_kernel<true>, jobs, mx);
}

void block_put_batched(Ctx& ctx, const std::vector<BlockPut>& jobs) {
  int64_t mx = 0;
  for (auto& j : jobs) mx = std::max<int64_t>(mx, int64_t(j.ra) * j.n * j.rb);
  launch_y(ctx, block_put_kernel, jobs, mx);
}

void kron_cores_batched(Ctx& ctx, const std::vector<KronCore>& jobs) {
  int64_t mx = 0, by = 0;
  for (auto& j : jobs) {
    const int64_t e = int64_t(j.RA) * j.n * j.RB;
    mx = std::max<int64_t>(mx, e);
    by += 8 * e;
  }
  ProfScope ps(ctx, "kron_core", 0.0, double(by));
  launch_y(ctx, kron_core_kernel, jobs, mx);
}

void mpo_apply_batched(Ctx& ctx, const std::vector<MpoCore>& jobs) {
  int64_t mx = 0;
  for (auto& j : jobs) mx = std::max<int64_t>(mx, int64_t(j.Ra) * j.rx * j.n * j.Ra2 * j.rx2);
  launch_y(ctx, mpo_kernel, jobs, mx);
}

namespace {
template <int MAXO, class TI, class TO>
void launch_op_apply(Ctx& ctx, const OpApply* d, int64_t njobs, int64_t bx) {
  TTR_LAUNCH(ctx, (op_apply_kernel<MAXO, TI, TO>), dim3(unsigned(bx), unsigned(njobs)), dim3(256), 0, d);
}
template <class TI, class TO>
void launch_op_apply_o(Ctx& ctx, int maxo, const OpApply* d, int64_t njobs, int64_t bx) {
  if (maxo <= 8) launch_op_apply<8, TI, TO>(ctx, d, njobs, bx);
  else if (maxo <= 16) launch_op_apply<16, TI, TO>(ctx, d, njobs, bx);
  else if (maxo <= 32) launch_op_apply<32, TI, TO>(ctx, d, njobs, bx);
  else launch_op_apply<64, TI, TO>(ctx, d, njobs, bx);
}
template <class TI>
void launch_op_apply_i(Ctx& ctx, Prec out, int maxo, const OpApply* d, int64_t njobs, int64_t bx) {
  if (out == Prec::F64) launch_op_apply_o<TI, double>(ctx, maxo, d, njobs, bx);
  else if (out == Prec::F32) launch_op_apply_o<TI, float>(ctx, maxo, d, njobs, bx);
  else launch_op_apply_o<TI, uint16_t>(ctx, maxo, d, njobs, bx);
}
}  // namespace

void op_apply_batched(Ctx& ctx, const std::vector<OpApply>& jobs, Prec in, Prec out) {
  if (jobs.empty()) return;
  int maxo = 1;
  int64_t mx = 1;
  double fl = 0.0, by = 0.0;
  for (const auto& j : jobs) {
    TTR_REQUIRE(j.n * j.rin <= 64 && j.n * j.rout <= 64, ST_INVALID_ARG,
                "implicit operator products need n*R <= 64 on every bond");
    maxo = std::max(maxo, j.n * j.rout);
    const int64_t nf = int64_t(j.X) * j.Y;
    mx = std::max(mx, nf);
    fl += 2.0 * nf * j.n * j.rin * j.n * j.rout;
    by += double(nf) * (j.n * j.rin * prec_bytes(in) + j.n * j.rout * prec_bytes(out));
  }
  ProfScope ps(ctx, "op_apply", fl, by);
  for (size_t s = 0; s < jobs.size(); s += 65535) {
    const size_t cnt = std::min<size_t>(65535, jobs.size() - s);
    const OpApply* d = ctx.upload(jobs.data() + s, cnt);
    const int64_t bx = std::min<int64_t>(1024, std::max<int64_t>(1, (mx + 255) / 256));
    if (in == Prec::F64) launch_op_apply_i<double>(ctx, out, maxo, d, int64_t(cnt), bx);
    else if (in == Prec::F32) launch_op_apply_i<float>(ctx, out, maxo, d, int64_t(cnt), bx);
    else launch_op_apply_i<uint16_t>(ctx, out, maxo, d, int64_t(cnt), bx);
    TTR_HIP(hipGetLastError());
  }
}

void pow2_normalize_batched(Ctx& ctx, const std::vector<Pow2Job>& jobs, Prec prec) {
  if (jobs.empty()) return;
  dim3 grid(unsigned(jobs.size())), block(256);
  if (arg_jobs_fit<SMALL_ARG_JOBS>(ctx, jobs)) {
    const auto a = arg_jobs<SMALL_ARG_JOBS>(jobs);
    if (prec == Prec::F64) TTR_LAUNCH(ctx, (pow2_kernel<double, true>), grid, block, 0, nullptr, a);
    else if (prec == Prec::F32) TTR_LAUNCH(ctx, (pow2_kernel<float, true>), grid, block, 0, nullptr, a);
    else TTR_LAUNCH(ctx, (pow2_kernel<uint16_t, true>), grid, block, 0, nullptr, a);
  } else {
    const Pow2Job* d = ctx.upload(jobs);
    const ArgJobs<Pow2Job, SMALL_ARG_JOBS, false> a{};
    if (prec == Prec::F64) TTR_LAUNCH(ctx, (pow2_kernel<double, false>), grid, block, 0, d, a);
    else if (prec == Prec::F32) TTR_LAUNCH(ctx, (pow2_kernel<float, false>), grid, block, 0, d, a);
    else TTR_LAUNCH(ctx, (pow2_kernel<uint16_t, false>), grid, block, 0, d, a);
  }
  TTR_HIP(hipGetLastError());
}

void sketch_generate(Ctx& ctx, const std::vector<SketchGen>& jobs, uint64_t seed, Prec prec) {
  if (jobs.empty()) return;
  int64_t mx = 0;
  for (auto& j : jobs) mx = std::max<int64_t>(mx, int64_t(j.n) * j.bvr * j.bvl * j.p);
  const SketchGen* d = ctx.upload(jobs);
  int64_t bx = std::min<int64_t>(1024, std::max<int64_t>(1, (mx + 255) / 256));
  TTR_LAUNCH(ctx, sketch_gen_kernel, dim3(unsigned(bx), unsigned(jobs.size())), dim3(256), 0, d, seed, int(prec));
  TTR_HIP(hipGetLastError());
}

void sketch_draws(Ctx& ctx, double* out, int64_t count, int core, uint64_t seed) {
  if (count <= 0) return;
  int64_t bx = std::min<int64_t>(4096, (count + 255) / 256);
  TTR_LAUNCH(ctx, sketch_draws_kernel, dim3(unsigned(bx)), dim3(256), 0, out, count, core, seed);
  TTR_HIP(hipGetLastError());
}

void randn_core(Ctx& ctx, double* out, int n, int r, int r2, int core, uint64_t seed) {
  const int64_t tot = int64_t(n) * r * r2;
  if (tot <= 0) return;
  int64_t bx = std::min<int64_t>(8192, (tot + 255) / 256);
  TTR_LAUNCH(ctx, randn_core_kernel, dim3(unsigned(bx)), dim3(256), 0, out, n, r, r2, core, seed);
  TTR_HIP(hipGetLastError());
}

void sketch_pack(Ctx& ctx, const double* ref, void* out, int n, int bvr, int bvl, int p, double scale, Prec prec) {
  const int64_t tot = int64_t(n) * bvr * bvl * p;
  if (tot <= 0) return;
  int64_t bx = std::min<int64_t>(4096, (tot + 255) / 256);
  TTR_LAUNCH(ctx, sketch_pack_kernel, dim3(unsigned(bx)), dim3(256), 0, ref, out, n, bvr, bvl, p, scale, int(prec));
  TTR_HIP(hipGetLastError());
}

void sum_arrays(Ctx& ctx, double* out, const std::vector<const double*>& ins, int64_t count) {
  if (count <= 0) return;
  const double* const* d = ctx.upload(ins.data(), ins.size());
  int64_t bx = std::min<int64_t>(4096, (count + 255) / 256);
  TTR_LAUNCH(ctx, sum_kernel, dim3(unsigned(bx)), dim3(256), 0, out, d, int(ins.size()), count);
  TTR_HIP(hipGetLastError());
}

void zero_fill(Ctx& ctx, void* p, size_t bytes) {
  // a fill kernel rather than hipMemsetAsync, so the call is safe inside captured graphs
  if (!bytes) return;
  TTR_REQUIRE((reinterpret_cast<uintptr_t>(p) & 3) == 0 && (bytes & 3) == 0, ST_INVALID_ARG, "zero_fill alignment");
  const int64_t words = int64_t(bytes / 4);
  int64_t bx = std::min<int64_t>(4096, (words + 255) / 256);
  TTR_LAUNCH(ctx, zero_kernel, dim3(unsigned(bx)), dim3(256), 0, static_cast<uint32_t*>(p), words);
  TTR_HIP(hipGetLastError());
}

}  // namespace ttr
