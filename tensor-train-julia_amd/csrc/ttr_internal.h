// Internal declarations of the MI355X TT-rounding library (libttround).
//
// Device layout of a TT core k: (r_{k-1}, n_k, r_k) column-major, i.e. element
// (a, i, b) at a + r_{k-1}*(i + n_k*b).  Both unfoldings are plain column-major
// matrices: the left unfolding (r_{k-1} n_k) x r_k and the right unfolding
// r_{k-1} x (n_k r_k), each with leading dimension r_{k-1}*n_k resp. r_{k-1}.
// Host-side (reference) cores are (n_k, r_{k-1}, r_k) column-major and are
// permuted on upload/download.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace ttr {

enum Status : int {
  ST_OK = 0,
  ST_INVALID_ARG = 1,
  ST_DIM_MISMATCH = 2,
  ST_OOM = 3,
  ST_NUMERICAL_FAILURE = 4,
  ST_DEVICE_ERROR = 5,
  ST_NOT_FOUND = 6,
};

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define TTR_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw ::ttr::Error(e_ == hipErrorOutOfMemory ? ::ttr::ST_OOM : ::ttr::ST_DEVICE_ERROR, \
                         std::string(#call) + ": " + hipGetErrorString(e_) + " (" __FILE__ \
                         ":" + std::to_string(__LINE__) + ")");                            \
  } while (0)

#define TTR_REQUIRE(cond, st, msg)                 \
  do {                                             \
    if (!(cond)) throw ::ttr::Error((st), (msg));  \
  } while (0)

enum class Prec : int { F64 = 0, F32 = 1, BF16 = 2 };

inline size_t prec_bytes(Prec p) { return p == Prec::F64 ? 8 : (p == Prec::F32 ? 4 : 2); }

// ---------------------------------------------------------------------------
// memory
// ---------------------------------------------------------------------------

// Bump arena of device memory, reset after every library call.  When one call
// needs more than the current block, further blocks are added and merged into
// a single block on reset, so steady-state calls allocate nothing.
// With a fixed region set (set_fixed), allocations come from that region only
// and a request beyond it throws: nothing is allocated (the entry points that
// must not allocate run this way).
class DevArena {
 public:
  ~DevArena();
  void* alloc(size_t bytes);
  template <class T>
  T* alloc_n(size_t n) { return static_cast<T*>(alloc(n * sizeof(T))); }
  void reset();  // caller must have synchronised the stream
  void release();
  void set_fixed(void* base, size_t cap) {
    fixed_ = static_cast<char*>(base);
    fixed_cap_ = cap;
    fixed_used_ = fixed_peak_ = 0;
  }
  void clear_fixed() { fixed_ = nullptr; fixed_cap_ = fixed_used_ = fixed_peak_ = 0; }
  bool fixed() const { return fixed_ != nullptr; }
  size_t fixed_free() const { return fixed_cap_ - fixed_used_; }
  // most bytes held at once since the last reset (the size a fixed region needs for the same call)
  size_t used() const;
  // A position in the arena; rewind(m) hands back everything allocated after it
  // (the blocks stay, later allocations of the call reuse them).
  struct Mark { size_t blk, off; };
  Mark mark() const;
  void rewind(const Mark& m);

 private:
  struct Block { char* p; size_t cap; size_t used; };
  size_t held() const;
  std::vector<Block> blocks_;
  size_t cur_ = 0, peak_ = 0;
  char* fixed_ = nullptr;
  size_t fixed_cap_ = 0, fixed_used_ = 0, fixed_peak_ = 0;
};

// Same scheme over pinned host memory (staging for descriptor arrays).
class PinnedArena {
 public:
  ~PinnedArena();
  void* alloc(size_t bytes);
  void reset();
  void release();
  void set_fixed(void* base, size_t cap) { fixed_ = static_cast<char*>(base); fixed_cap_ = cap; fixed_used_ = 0; }
  void clear_fixed() { fixed_ = nullptr; fixed_cap_ = fixed_used_ = 0; }
  size_t used() const;

 private:
  struct Block { char* p; size_t cap; size_t used; };
  std::vector<Block> blocks_;
  char* fixed_ = nullptr;
  size_t fixed_cap_ = 0, fixed_used_ = 0;
};

// Caching allocator for long-lived objects (TT storage): frees go to per-size
// free lists and are reused, so repeated calls of the same shape do not call
// hipMalloc.
class CachingAlloc {
 public:
  ~CachingAlloc();
  void* alloc(size_t bytes);
  void free(void* p);
  void release_cached();

 private:
  static size_t round_size(size_t b);
  std::multimap<size_t, void*> free_;
  std::unordered_map<void*, size_t> live_;
};

// Per-kernel / per-phase timing with hipEvents on the handle's stream
// (enabled by ttr_profile_enable).  Each record keeps the algorithmic flops and
// bytes of the launch so achieved rates can be computed without counters.
struct ProfStat {
  double ms = 0.0;
  int64_t launches = 0;
  double flops = 0.0;
  double bytes = 0.0;
};

class Profiler {
 public:
  ~Profiler();
  bool on = false;
  // 1: every timed launch; 2: only the GEMM classes and the sweep phases (two
  // events per launch cost ~3 us of host time each, so a run that needs only the
  // dominant kernel's times leaves the others out)
  int level = 1;
  bool records(const char* name) const {
    return level == 1 || std::strncmp(name, "gemm_", 5) == 0 || std::strncmp(name, "phase:", 6) == 0;
  }
  // returns a token; call stop(token) after enqueuing the launch
  // name2 (optional): the record is also added to that second statistic
  int start(hipStream_t s, const char* name, double flops, double bytes, const char* name2 = nullptr);
  void stop(hipStream_t s, int token);
  // adds every closed record to stats (the stream must be synchronised); open
  // records (start without stop yet) stay pending
  void collect();
  void reset() { stats.clear(); }
  std::map<std::string, ProfStat> stats;

 private:
  struct Pending { hipEvent_t a, b; std::string name, name2; double flops, bytes; bool stopped; };
  std::map<int, Pending> pending_;
  int next_ = 0, stopped_ = 0;
  std::vector<hipEvent_t> pool_;
  hipEvent_t get_event();
};

// HIP graphs of whole calls: Capture records the launches and descriptor
// uploads of a call into a graph; Replay reruns the host logic of a call of the
// same shape, writing its descriptors into the same pinned staging (the graph's
// copy nodes read them) without launching anything, then launches the graph.
enum class GraphMode : int { Off = 0, Capture = 1, Replay = 2 };

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  DevArena ws;
  PinnedArena pin;
  CachingAlloc cache;
  Profiler prof;
  // GEMM tile choice per (call site, precisions, size class), measured on first use
  std::unordered_map<std::string, int> gemm_cfg;
  bool gemm_autotune = true;
  // GEMM tile configuration forced for every call (-1: choose per site)
  int gemm_force_cfg = -1;
  // one line per GEMM dispatch, in submission order (TTR_LAUNCH_LOG=path; null: off), so
  // per-dispatch profiler counters can be attributed to call sites
  FILE* launch_log = nullptr;
  // skip the SVD of a bond whose singular values a Cholesky certificate shows
  // the floor rule keeps (see chol_certify_batched)
  bool svd_certify = true;
  // Jacobi sweeps after which a matrix that still rotates is a numerical failure
  int jacobi_max_sweeps = 40;
  // Jacobi rotation threshold: |G_ij| > scale * max(p, 16) eps sqrt(G_ii G_jj)
  double jacobi_tol_scale = 1.0;
  // sort the columns by decreasing norm before every Jacobi sweep (de Rijk)
  bool jacobi_sort = true;
  // QR batches whose matrices have at most QR_SMALL_MAX_M rows and whose total
  // size is at most this many bytes run one workgroup per matrix (qr_small_batched);
  // larger batches use the blocked path whose updates spread over the chip
  size_t qr_small_bytes = size_t(96) << 20;
  // 0: the QR path by size and batch; 1: never the one-launch QR (blocked path); 2: the
  // blocked path with streaming panels of 32 at every height; 3: the one-launch QR
  // for every batch whose matrices have at most QR_SMALL_MAX_M rows
  int qr_path = 0;
  // 64-column blocks per compact WY group of the blocked QR (1, 2 or 4).  Measured on
  // the C3 workload: 2 (2.13 s per step) beats 4 (2.17 s: faster trailing updates, but
  // the wider group's V carries more zeros into the formation of Q, and the in-group
  // updates and T merges grow)
  int qr_group = 2;
  // partial sums of split-K GEMMs, reused by every launch on the stream
  double* splitk = nullptr;
  size_t splitk_elems = 0;
  double* fsplit = nullptr;  // the same inside the fixed workspace region
  size_t fsplit_elems = 0;
  // with the workspace in fixed mode: taken from it, or nullptr when it does
  // not fit (the GEMM then runs without splitting K)
  double* splitk_buffer(size_t elems);
  size_t splitk_peak = 0;  // largest request of the current call (elements)
  GraphMode gmode = GraphMode::Off;
  // every launch uploads its descriptors (no descriptors in kernel arguments): set
  // while a call measures the workspace that its graph capture will need
  bool upload_all = false;
  // descriptors of a launch may travel in its kernel arguments
  bool args_ok() const { return gmode == GraphMode::Off && !upload_all; }
  bool launching() const { return gmode != GraphMode::Replay; }
  // host-side record of a captured or replayed call: sizes of every upload and
  // every launch, compared between the capture and each replay
  std::vector<uint64_t> gsig;
  // copies a host array into fresh device workspace (asynchronously, through pinned memory)
  template <class T>
  T* upload(const T* host, size_t n) {
    if (n == 0) return nullptr;
    void* h = pin.alloc(n * sizeof(T));
    std::memcpy(h, host, n * sizeof(T));
    T* d = ws.alloc_n<T>(n);
    if (gmode != GraphMode::Off) {
      gsig.push_back(uint64_t(n * sizeof(T)));
      gsig.push_back(uint64_t(reinterpret_cast<uintptr_t>(h)) ^ (uint64_t(reinterpret_cast<uintptr_t>(d)) << 1));
    }
    if (launching()) TTR_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, stream));
    return d;
  }
  template <class T>
  T* upload(const std::vector<T>& v) { return upload(v.data(), v.size()); }
  void sync() { TTR_HIP(hipStreamSynchronize(stream)); }
  void end_call() { sync(); if (prof.on) prof.collect(); ws.reset(); pin.reset(); }
};

// Device workspace taken inside the scope goes back to the arena when the scope
// ends, so a loop of calls (a QR per core step) reuses one region instead of
// growing the arena by each call's share.  All device work runs on ctx.stream,
// so the kernels that read the region are queued ahead of any that overwrite it.
// Nothing allocated inside the scope may be used after it; a split-K buffer made
// inside it is dropped too.
class WsScope {
 public:
  explicit WsScope(Ctx& c) : c_(c), m_(c.ws.mark()), fs_(c.fsplit), fse_(c.fsplit_elems) {}
  ~WsScope() {
    c_.ws.rewind(m_);
    c_.fsplit = fs_;
    c_.fsplit_elems = fse_;
  }
  WsScope(const WsScope&) = delete;
  WsScope& operator=(const WsScope&) = delete;

 private:
  Ctx& c_;
  DevArena::Mark m_;
  double* fs_;
  size_t fse_;
};

// FNV-1a hash of a launch (kernel, grid, block, LDS bytes, argument bytes), kept
// in ctx.gsig while a call is captured or replayed: a replay whose launches
// differ in any of these from the captured ones is not run from the graph
template <class... A>
inline void graph_note_launch(Ctx& ctx, const void* kernel, dim3 grid, dim3 block, size_t shmem, const A&... a) {
  uint64_t x = 0xcbf29ce484222325ull;
  auto mix = [&x](const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) x = (x ^ c[i]) * 0x100000001b3ull;
  };
  mix(&kernel, sizeof(kernel));
  mix(&grid, sizeof(grid));
  mix(&block, sizeof(block));
  mix(&shmem, sizeof(shmem));
  (mix(&a, sizeof(A)), ...);
  ctx.gsig.push_back(x);
}

// Kernel launch on ctx.stream.  During graph capture or replay the launch is
// recorded in ctx.gsig; in replay it is not enqueued (the graph holds it).
#define TTR_LAUNCH(ctx, kernel, grid, block, shmem, ...)                                                   \
  do {                                                                                                     \
    if ((ctx).gmode != ::ttr::GraphMode::Off)                                                              \
      ::ttr::graph_note_launch((ctx), reinterpret_cast<const void*>(kernel), dim3(grid), dim3(block),      \
                               size_t(shmem), __VA_ARGS__);                                                \
    if ((ctx).launching()) {                                                                               \
      hipLaunchKernelGGL(kernel, grid, block, shmem, (ctx).stream, __VA_ARGS__);                           \
      TTR_HIP(hipGetLastError());                                                                          \
    }                                                                                                      \
  } while (0)

// Job descriptors of a launch with at most N jobs, passed in the kernel
// arguments instead of an upload: no copy dispatch and no dependent launch
// boundary ahead of the kernel.  Launches with more jobs, and launches recorded
// into a graph (whose replays rewrite uploaded descriptors in place), pass a
// device array; the kernel then gets the empty ArgJobs<T, N, false>.
template <class T, int N, bool INL> struct ArgJobs {};
template <class T, int N> struct ArgJobs<T, N, true> { T v[N]; };
template <int N, class T>
inline bool arg_jobs_fit(const Ctx& ctx, const std::vector<T>& jobs) {
  return ctx.args_ok() && !jobs.empty() && jobs.size() <= size_t(N);
}
template <int N, class T>
inline ArgJobs<T, N, true> arg_jobs(const std::vector<T>& jobs) {
  ArgJobs<T, N, true> a{};
  std::copy(jobs.begin(), jobs.end(), a.v);
  return a;
}

// RAII timing scope around one or more launches on ctx.stream (off inside graphs)
struct ProfScope {
  Ctx& ctx;
  int tok;
  ProfScope(Ctx& c, const char* name, double flops = 0.0, double bytes = 0.0, const char* name2 = nullptr)
      : ctx(c),
        tok(c.prof.on && name && c.gmode == GraphMode::Off && c.prof.records(name)
                ? c.prof.start(c.stream, name, flops, bytes, name2)
                : -1) {}
  ~ProfScope() { if (tok >= 0) ctx.prof.stop(ctx.stream, tok); }
};

// ---------------------------------------------------------------------------
// grouped GEMM: C_p = alpha_p * sum_s scale_s A_s B_s + beta_p C_p
// Element (i,l) of A_s is at A_s[i*a_rs + l*a_cs] (same for B and C), so any
// transpose is expressed through the strides.
// ---------------------------------------------------------------------------

enum AMode : int32_t {
  AM_PLAIN = 0,
  AM_VT_UNIT = 1,  // A(i,l) = V(l,i) with V unit lower trapezoidal (zero if l<i, one if l==i)
  AM_V_UNIT = 2,   // A(i,l) = V(i,l) unit lower trapezoidal (zero if i<l, one if i==l)
};

struct GemmSeg {
  const void* A;
  const void* B;
  int64_t a_rs, a_cs, b_rs, b_cs;
  int32_t K;
  int32_t a_mode;
  double scale;
  const int32_t* exp;  // optional: scale *= 2^(exp[0] - max over the problem's segments)
};

struct GemmProb {
  void* C;
  int64_t c_rs, c_cs;
  int32_t M, N;
  int32_t seg0, nseg;
  double alpha, beta;
  int32_t tile0, tiles_n, tiles_mn;
  // split K: ksplit chunks of kc k-tiles each (the k-tiles of all segments in
  // order); chunk c writes its raw sum to part + c*M*N (i fastest) and a
  // reduction adds the chunks in order, then applies alpha and beta
  int32_t ksplit, kc;
  double* part;
};

class GemmBatch {
 public:
  // starts a problem; segments are added with seg()
  void prob(void* C, int64_t c_rs, int64_t c_cs, int M, int N, double alpha, double beta);
  void seg(const void* A, int64_t a_rs, int64_t a_cs, const void* B, int64_t b_rs, int64_t b_cs,
           int K, double scale = 1.0, int a_mode = AM_PLAIN, const int32_t* exp = nullptr);
  // convenience: single-segment problem
  void add(int M, int N, int K, double alpha, const void* A, int64_t a_rs, int64_t a_cs,
           const void* B, int64_t b_rs, int64_t b_cs, double beta, void* C, int64_t c_rs,
           int64_t c_cs, int a_mode = AM_PLAIN) {
    prob(C, c_rs, c_cs, M, N, alpha, beta);
    seg(A, a_rs, a_cs, B, b_rs, b_cs, K, 1.0, a_mode);
  }
  bool empty() const { return probs_.empty(); }
  void clear() { probs_.clear(); segs_.clear(); }
  std::vector<GemmProb> probs_;
  std::vector<GemmSeg> segs_;
};

// Launch all problems of the batch in one grouped launch.  ta/tb/tc are the
// storage precisions of A, B, C; math is the MFMA precision (F64, F32 or BF16,
// the latter two accumulating in FP32).
// site: optional call-site label; with profiling on, the launch is also
// recorded under "site:<label>"
void gemm_run(Ctx& ctx, GemmBatch& b, Prec ta, Prec tb, Prec tc, Prec math, const char* site = nullptr);

// ---------------------------------------------------------------------------
// Householder QR (LAPACK dgeqrf/dorgqr conventions), batched
// ---------------------------------------------------------------------------
struct QRJob {
  double* A;       // m x n column-major, overwritten by R (upper) and V (strictly lower)
  int64_t lda;
  int m, n;
  double* Q;       // optional output: thin Q, m x min(m,n)
  int64_t ldq;
};
// factorizes every job; forms thin Q into job.Q when it is non-null
void qr_batched(Ctx& ctx, std::vector<QRJob>& jobs);
// the same in one launch, one workgroup per matrix (m <= QR_SMALL_MAX_M); used by
// qr_batched for batches whose matrices all fit (see Ctx::qr_small_bytes)
constexpr int QR_SMALL_MAX_M = 1024;
// and only for batches of at least this many matrices (one per CU), or of single panels
constexpr size_t QR_SMALL_MIN_BATCH = 256;
void qr_small_batched(Ctx& ctx, const std::vector<QRJob>& jobs);

// ---------------------------------------------------------------------------
// one-sided Jacobi SVD of square matrices, batched
// ---------------------------------------------------------------------------
struct SVDJob {
  double* B;   // p x p (ld ldb): overwritten by B*J (columns U_j s_j)
  int64_t ldb;
  double* J;   // p x p (ld ldj): accumulated rotations (right singular vectors)
  int64_t ldj;
  double* sv;  // p singular values (column norms of B*J), unsorted
  int p;
  int32_t* info;  // optional, 8 counters zeroed by the caller: [0] sweeps used, [1..3] wall-clock
                  // ticks of the Gram / inner / apply phases summed over visits, [4] visits,
                  // [5] visits that applied a rotation, [6..7] unused
};
// returns the number of columns per visit (for the work counters)
int jacobi_svd_batched(Ctx& ctx, std::vector<SVDJob>& jobs);
// Certificate that a truncated SVD of B keeps all its singular values: attempts
// the Cholesky factorisation of G - tau' I for the Gram G = B^T B (m x m, ld,
// overwritten), tau' = tol^2 trace(G) (1 + 2^-6) + 8 (m + 1) 2^-53 trace(G).
// *flag = 1 when it succeeds (every singular value s of B has
// s^2 > tol^2 sum s^2 with margin), 0 otherwise (nothing is decided).
constexpr int CHOL_MAX_M = 1024;
struct CholJob {
  double* G;
  int64_t ld;
  int m;
  double tol;
  int32_t* flag;  // device
};
void chol_certify_batched(Ctx& ctx, const std::vector<CholJob>& jobs);
// pseudo-inverse weights after a Jacobi SVD: column c of BJ (p x p, ld) is scaled
// by 1/sv_c^2 when sv_c > eps * max(sv), else zeroed
struct PinvWeights {
  double* BJ;
  int64_t ld;
  const double* sv;
  int p;
  double eps;
};
void pinv_weights_batched(Ctx& ctx, const std::vector<PinvWeights>& jobs);

// ---------------------------------------------------------------------------
// elementwise / data-movement kernels
// ---------------------------------------------------------------------------
struct Copy2D {  // out[b*obs + i*ors + j*ocs] = alpha * in[b*ibs + i*irs + j*ics]
  const double* in;
  double* out;
  int64_t irs, ics, ibs, ors, ocs, obs;
  int32_t M, N, nb;
  double alpha;
  int32_t tri;  // 0: full copy; 1: keep i <= j (upper), zero below; 2: keep i >= j (lower), zero above;
                // 3: unit lower (keep i > j, one on the diagonal, zero above)
};
inline Copy2D copy_job(const double* in, int64_t irs, int64_t ics, double* out, int64_t ors, int64_t ocs, int M, int N,
                       double alpha = 1.0, int tri = 0) {
  Copy2D c{};
  c.in = in; c.out = out; c.irs = irs; c.ics = ics; c.ors = ors; c.ocs = ocs;
  c.ibs = 0; c.obs = 0; c.M = M; c.N = N; c.nb = 1; c.alpha = alpha; c.tri = tri;
  return c;
}
void copy2d_batched(Ctx& ctx, const std::vector<Copy2D>& jobs);

struct Fill2D {  // out[i + j*ld] = (i == j) ? diag : 0   for i < M, j < N
  double* out;
  int64_t ld;
  int32_t M, N;
  double diag;
};
void fill2d_batched(Ctx& ctx, const std::vector<Fill2D>& jobs);

// gathers columns: out[:, c] = in[:, idx[c]] (M rows, ncol columns, optional scale per column)
struct GatherCols {
  const double* in;
  double* out;
  int64_t ldi, ldo;
  const int32_t* idx;  // device
  int32_t M, ncol;
};
void gather_cols_batched(Ctx& ctx, const std::vector<GatherCols>& jobs);

// power-of-two normalisation of a matrix in place (storage precision prec):
// e = exponent of max |x|; x *= 2^-e; exp_out = exp_in + e
struct Pow2Job {
  void* X;
  int64_t ld;
  int32_t M, N;
  const int32_t* exp_in;  // may be null (treated as 0)
  int32_t* exp_out;
};
void pow2_normalize_batched(Ctx& ctx, const std::vector<Pow2Job>& jobs, Prec prec);

// sketch draws: Philox4x32-10 + Box-Muller (FP64), see rng_normal in kernels.hip
struct SketchGen {
  void* out;          // destination (precision prec), layout (n, bvr, bvl, p) with n fastest
  double* ref_out;    // optional FP64 copy in the reference order (bvr*n, bvl, p), times ref_scale
  int32_t n, bvr, bvl, p;  // bvr = left rank of the block, bvl = right rank
  int32_t core;       // core index k (0-based), part of the counter
  double scale;
  double ref_scale = 1.0;
};
void sketch_generate(Ctx& ctx, const std::vector<SketchGen>& jobs, uint64_t seed, Prec prec);
// raw draws in reference order for a given core (used by tests)
void sketch_draws(Ctx& ctx, double* out, int64_t count, int core, uint64_t seed);
// N(0,1) TT core (rand_tt semantics): draw e in reference order (i + n*(a + r*b)) of
// core k goes to device position a + r*(i + n*b); counter stream distinct from sketches
void randn_core(Ctx& ctx, double* out, int n, int r, int r2, int core, uint64_t seed);
// converts FP64 blocks in reference order (bvr*n, bvl, p) to the device layout
void sketch_pack(Ctx& ctx, const double* ref, void* out, int n, int bvr, int bvl, int p, double scale,
                 Prec prec);

// sum of nterm equally shaped arrays: out[e] = sum_t in_t[e]
void sum_arrays(Ctx& ctx, double* out, const std::vector<const double*>& ins, int64_t count);

// block-diagonal placement of a core into a larger core (TT sum)
struct BlockPut {
  const double* in;  // (ra, n, rb) device layout
  double* out;       // (RA, n, RB) device layout
  int32_t ra, rb, n, RA, RB, a0, b0;
  double scale;
};
void block_put_batched(Ctx& ctx, const std::vector<BlockPut>& jobs);

// MPO x TT core: y[(al,nu), i, (al',nu')] = sum_j A[i, j, al, al'] x[nu, j, nu']
struct MpoCore {
  const double* A;  // (n, n, Ra, Ra') column-major (reference layout)
  const double* x;  // (rx, n, rx') device layout
  double* y;        // (Ra*rx, n, Ra'*rx') device layout
  int32_t n, Ra, Ra2, rx, rx2;
};
void mpo_apply_batched(Ctx& ctx, const std::vector<MpoCore>& jobs);

// Entrywise (Hadamard) product of m TTs, one core: out[(a, i, b)] =
// prod_f in_f[(a_f, i, b_f)] with a = a_0 + ra_0 (a_1 + ra_1 (...)) (factor 0
// fastest), likewise b; every index below 2^31
constexpr int KRON_MAX = 8;
struct KronCore {
  const double* in[KRON_MAX];  // (ra_f, n, rb_f) device layout
  double* out;                 // (RA, n, RB) device layout
  int32_t ra[KRON_MAX], rb[KRON_MAX];
  int32_t m, n, RA, RB;
};
void kron_cores_batched(Ctx& ctx, const std::vector<KronCore>& jobs);

// Local action of an operator core on the (mode, bond) pairs of every fiber
// (x, y) of a 2-D grid, without forming an MPO x TT product:
//   out[(om, ob)] = scale * sum_{im, ib} H(om, ob, im, ib) in[(im, ib)],
// H(om, ob, im, ib) = H[om*h_om + ob*h_ob + im*h_im + ib*h_ib] (H is a core of a
// TToperator in the reference layout), in[(im, ib)] at in + x*is_x + y*is_y +
// im*is_m + ib*is_b, out likewise.  n*rin and n*rout must not exceed 64.
struct OpApply {
  const void* in;
  void* out;
  const double* H;
  int64_t h_om, h_ob, h_im, h_ib;
  int32_t n, rin, rout, X, Y;
  int64_t is_m, is_b, is_x, is_y;
  int64_t os_m, os_b, os_x, os_y;
  double scale;
};
void op_apply_batched(Ctx& ctx, const std::vector<OpApply>& jobs, Prec in, Prec out);

void zero_fill(Ctx& ctx, void* p, size_t bytes);

}  // namespace ttr
