// libttround: TT rounding drivers and the C-ABI (include/ttround.h).
//
// Every driver works on a batch of TTs in lock-step over the core index, so a
// sweep step over the whole batch is a handful of grouped launches (GEMM, QR
// panels, Jacobi, copies) on the handle's stream.
#include <algorithm>
#include <array>
#include <cmath>
#include <mutex>
#include <numeric>

#include "../../include/ttround.h"
#include <cstdlib>

#include "ttr_internal.h"

using namespace ttr;

// Device scratch of the entry points that allocate nothing (ttr_dot_noalloc):
// environments, temporaries, split-K partial sums and GEMM descriptors of one
// call.  Each GPU the library runs on holds its own copy until the process
// ends; every such call on a GPU uses that GPU's copy, so calls that overlap on
// one GPU (two handles on two host threads) race.  Host code never passes the
// array's name to a kernel: it uses the address hipGetSymbolAddress returned
// for the device (scratch_base), taken once per device in ttr_create.
constexpr size_t TTR_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) char scratch[TTR_SCRATCH_BYTES];
// pinned host staging of the descriptors, allocated per handle in ttr_create
constexpr size_t TTR_STAGE_BYTES = size_t(4) << 20;

namespace {

std::mutex scratch_mu;
std::vector<void*> scratch_base;  // per device: the address of `scratch` there

// the address of this device's scratch (the device must be current)
void* scratch_for_device(int device) {
  std::lock_guard<std::mutex> g(scratch_mu);
  if (scratch_base.size() <= size_t(device)) scratch_base.resize(size_t(device) + 1, nullptr);
  if (!scratch_base[size_t(device)]) {
    void* p = nullptr;
    TTR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TTR_REQUIRE(p != nullptr, ST_DEVICE_ERROR, "hipGetSymbolAddress returned no address for the scratch array");
    scratch_base[size_t(device)] = p;
  }
  return scratch_base[size_t(device)];
}

struct DevTT {
  int d = 0;
  std::vector<int64_t> dims, rks, ot;
  std::vector<double*> cores;
  std::vector<void*> allocs;
  int64_t core_elems(int k) const { return rks[k] * dims[k] * rks[k + 1]; }
};

struct DevTTO {
  int d = 0;
  std::vector<int64_t> dims, rks;
  std::vector<double*> cores;  // reference layout (n, n, R, R')
  // cores with n*max(R, R') > 64 also as the matrix Hm[(i + n b), (beta + R j)]
  // = H[i, j, beta, b] ((n R') x (R n), column-major), for GEMM contractions of
  // bonds too wide for the LDS-staged local action; nullptr elsewhere
  std::vector<double*> hm;
  void* base = nullptr;
  void* base_m = nullptr;
  DevTTO() = default;
  DevTTO(const DevTTO&) = delete;
  DevTTO& operator=(const DevTTO&) = delete;
  ~DevTTO() {
    if (base) (void)hipFree(base);
    if (base_m) (void)hipFree(base_m);
  }
  bool wide(int k) const { return dims[k] * std::max(rks[k], rks[k + 1]) > 64; }
};

int64_t wrap_mul(int64_t a, int64_t b) { return int64_t(uint64_t(a) * uint64_t(b)); }

int64_t wrap_prod(const std::vector<int64_t>& v, size_t lo, size_t hi) {
  int64_t p = 1;
  for (size_t i = lo; i < hi; ++i) p = wrap_mul(p, v[i]);
  return p;
}

// r_and_d_to_rks (src/tt_rounding.jl:6-24), including its Int64 overflow guards
std::vector<int64_t> r_and_d_to_rks(const std::vector<int64_t>& rks, const std::vector<int64_t>& dims, int64_t rmax) {
  const size_t d = dims.size();
  std::vector<int64_t> out(rks.size(), 1);
  for (size_t i = 0; i < d; ++i) {
    const int64_t right = wrap_prod(dims, i, d), left = wrap_prod(dims, 0, i);
    int64_t v = std::min(rks[i], rmax);
    if (right > 0) v = std::min(v, right);
    if (left > 0) v = std::min(v, left);
    out[i] = v;
  }
  return out;
}

// floor (src/tt_rounding.jl:297-315) on descending singular values
int floor_rule(const std::vector<double>& s, double tol, double degen_tol = 1e-5) {
  const int d = int(s.size());
  if (tol == 0.0) return d;
  int i = d;
  double weight = 0.0, norm2 = 0.0;
  for (double v : s) norm2 += v * v;
  const double thr = tol * tol * norm2;
  while (i > 0 && weight < thr) {
    weight += s[i - 1] * s[i - 1];
    --i;
  }
  ++i;
  while (i < d) {
    const double ratio = s[i] / s[i - 1];
    if (std::isfinite(ratio) && std::fabs(1.0 - ratio) <= degen_tol) ++i;
    else break;
  }
  return i;
}

Prec to_prec(int p) {
  TTR_REQUIRE(p >= 0 && p <= 2, ST_INVALID_ARG, "sketch_prec must be 0 (f64), 1 (f32) or 2 (bf16)");
  return Prec(p);
}

float bf16_bits_to_float(uint16_t h) {
  uint32_t u = uint32_t(h) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

}  // namespace

struct ttr_handle_s {
  Ctx ctx;
  std::string err;
  int64_t next_id = 1;
  std::unordered_map<int64_t, std::unique_ptr<DevTT>> tts;
  std::unordered_map<int64_t, std::unique_ptr<DevTTO>> ops;
  double* one = nullptr;  // device constant 1.0
  void* scratch = nullptr;  // this device's scratch array (set in ttr_create)
  void* stage = nullptr;    // pinned host staging of the non-allocating entry points

  // HIP graphs of ttr_round_rand, one per call shape (see round_rand_graphed)
  struct GraphEntry {
    int state = 0;            // 0: run once eagerly, 1: captured, -1: never graphed
    size_t ws_need = 0, pin_need = 0;
    size_t split_elems = 0;   // split-K partial sums: the first call's largest request
    void* dev = nullptr;      // fixed device workspace of the graph
    void* host = nullptr;     // fixed pinned staging its copy nodes read
    size_t dev_cap = 0, host_cap = 0;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> sig;  // uploads and launches recorded at capture
    uint64_t last_use = 0;
  };
  bool graphs = false;
  std::map<std::string, GraphEntry> gcache;
  uint64_t gclock = 0;
  int64_t g_captures = 0, g_replays = 0, g_fallbacks = 0;
  static void graph_free(GraphEntry& e) {
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    if (e.dev) (void)hipFree(e.dev);
    if (e.host) (void)hipHostFree(e.host);
    e.exec = nullptr;
    e.dev = e.host = nullptr;
    e.dev_cap = e.host_cap = 0;
  }
  void graph_clear() {
    (void)hipStreamSynchronize(ctx.stream);
    for (auto& kv : gcache) graph_free(kv.second);
    gcache.clear();
  }

  DevTT* get(int64_t id) {
    auto it = tts.find(id);
    TTR_REQUIRE(it != tts.end(), ST_NOT_FOUND, "unknown TT id " + std::to_string(id));
    return it->second.get();
  }
  DevTTO* get_op(int64_t id) {
    auto it = ops.find(id);
    TTR_REQUIRE(it != ops.end(), ST_NOT_FOUND, "unknown TT operator id " + std::to_string(id));
    return it->second.get();
  }
  // new TT with all cores allocated (contents undefined)
  std::unique_ptr<DevTT> make_tt(const std::vector<int64_t>& dims, const std::vector<int64_t>& rks,
                                 const std::vector<int64_t>& ot) {
    auto t = std::make_unique<DevTT>();
    t->d = int(dims.size());
    t->dims = dims;
    t->rks = rks;
    t->ot = ot;
    int64_t total = 0;
    for (int k = 0; k < t->d; ++k) total += t->core_elems(k);
    char* base = static_cast<char*>(ctx.cache.alloc(size_t(std::max<int64_t>(total, 1)) * 8));
    t->allocs.push_back(base);
    int64_t off = 0;
    for (int k = 0; k < t->d; ++k) {
      t->cores.push_back(reinterpret_cast<double*>(base) + off);
      off += t->core_elems(k);
    }
    return t;
  }
  // TT whose cores are filled in later, one cache allocation per core
  std::unique_ptr<DevTT> make_tt_lazy(const std::vector<int64_t>& dims) {
    auto t = std::make_unique<DevTT>();
    t->d = int(dims.size());
    t->dims = dims;
    t->rks.assign(dims.size() + 1, 1);
    t->ot.assign(dims.size(), 0);
    t->cores.assign(dims.size(), nullptr);
    return t;
  }
  double* core_alloc(DevTT* t, int64_t elems) {
    void* p = ctx.cache.alloc(size_t(std::max<int64_t>(elems, 1)) * 8);
    t->allocs.push_back(p);
    return static_cast<double*>(p);
  }
  int64_t add(std::unique_ptr<DevTT> t) {
    const int64_t id = next_id++;
    tts[id] = std::move(t);
    return id;
  }
  void free_tt(DevTT* t) {
    for (void* p : t->allocs) ctx.cache.free(p);
    t->allocs.clear();
  }
};

namespace {

// ------------------------------------------------------------------ helpers

void zeros_like_ranks(ttr_handle_s* h, DevTT* t) {
  for (int k = 0; k < t->d; ++k) zero_fill(h->ctx, t->cores[k], size_t(t->core_elems(k)) * 8);
}

void check_dims_equal(const DevTT* a, const DevTT* b) {
  TTR_REQUIRE(a->dims == b->dims, ST_DIM_MISMATCH, "Incompatible dimensions");
}

bool is_leftorth(const DevTT* x) {
  for (int k = 0; k + 1 < x->d; ++k)
    if (x->ot[k] != 1) return false;
  return true;
}
bool is_rightorth(const DevTT* x) {
  for (int k = 1; k < x->d; ++k)
    if (x->ot[k] != -1) return false;
  return true;
}

std::vector<double> frob_norms_host(ttr_handle_s* h, const std::vector<const double*>& ptrs,
                                    const std::vector<int64_t>& counts) {
  // |x|_F through a one-column GEMM: s = x^T x
  Ctx& ctx = h->ctx;
  double* out = ctx.ws.alloc_n<double>(std::max<size_t>(1, ptrs.size()));
  GemmBatch gb;
  for (size_t i = 0; i < ptrs.size(); ++i) {
    if (counts[i] == 0) {
      gb.prob(out + i, 1, 1, 1, 1, 0.0, 0.0);
      gb.seg(h->one, 1, 1, h->one, 1, 1, 1, 1.0);
      continue;
    }
    gb.add(1, 1, int(counts[i]), 1.0, ptrs[i], 1, 1, ptrs[i], 1, 1, 0.0, out + i, 1, 1);
  }
  gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "norm");
  std::vector<double> res(ptrs.size());
  ctx.sync();
  TTR_HIP(hipMemcpy(res.data(), out, res.size() * 8, hipMemcpyDeviceToHost));
  for (auto& v : res) v = std::sqrt(std::max(v, 0.0));
  return res;
}

// ------------------------------------------------------------ orthogonalize
// orthogonalize(x; i=root) (src/tt_rounding.jl:65-97) for a batch
std::vector<std::unique_ptr<DevTT>> orthogonalize_batch(ttr_handle_s* h, const std::vector<DevTT*>& xs, int root) {
  Ctx& ctx = h->ctx;
  const size_t nb = xs.size();
  std::vector<std::unique_ptr<DevTT>> outs(nb);
  std::vector<std::vector<int64_t>> yr(nb);
  int maxd = 0;
  for (size_t q = 0; q < nb; ++q) {
    const DevTT* x = xs[q];
    TTR_REQUIRE(root >= 1 && root <= x->d, ST_DIM_MISMATCH, "Impossible orthogonalization");
    const int d = x->d, i0 = root - 1;
    std::vector<int64_t> r = r_and_d_to_rks(x->rks, x->dims, 1024);
    for (int j = 0; j < i0; ++j) r[j + 1] = std::min(x->dims[j] * r[j], x->rks[j + 1]);
    for (int j = d - 1; j > i0; --j) r[j] = std::min(x->dims[j] * r[j + 1], x->rks[j]);
    std::vector<int64_t> ot(d, 0);
    for (int j = 0; j < i0; ++j) ot[j] = 1;
    for (int j = i0 + 1; j < d; ++j) ot[j] = -1;
    outs[q] = h->make_tt(x->dims, r, ot);
    yr[q] = r;
    maxd = std::max(maxd, d);
  }
  // workspace: FR / FL factors and a working matrix per item
  std::vector<double*> FR(nb), FL(nb), Abuf(nb), Qw(nb);
  std::vector<Fill2D> fz;
  for (size_t q = 0; q < nb; ++q) {
    const DevTT* x = xs[q];
    int64_t mx_a = 1, mx_f = 1;
    for (int j = 0; j < x->d; ++j) {
      mx_a = std::max(mx_a, std::max(yr[q][j], x->rks[j]) * x->dims[j] * std::max(yr[q][j + 1], x->rks[j + 1]));
      mx_f = std::max(mx_f, std::max(yr[q][j], x->rks[j]) * std::max(yr[q][j], x->rks[j]));
    }
    FR[q] = ctx.ws.alloc_n<double>(size_t(mx_f));
    FL[q] = ctx.ws.alloc_n<double>(size_t(mx_f));
    Abuf[q] = ctx.ws.alloc_n<double>(size_t(mx_a));
    Qw[q] = ctx.ws.alloc_n<double>(size_t(mx_a));
    fz.push_back(Fill2D{FR[q], 1, 1, 1, 0.0});
    fz.push_back(Fill2D{FL[q], 1, 1, 1, 0.0});
  }
  fill2d_batched(ctx, fz);
  GemmBatch gb;
  // ---- left-to-right (cores 0..root-2): A = FR X_j, QR, core = Q, FR = R
  for (int j = 0; j < root - 1; ++j) {
    std::vector<QRJob> qj;
    std::vector<Copy2D> cp;
    for (size_t q = 0; q < nb; ++q) {
      const DevTT* x = xs[q];
      const int64_t n = x->dims[j], rp = yr[q][j], R0 = x->rks[j], R1 = x->rks[j + 1];
      if (j == 0)
        cp.push_back(copy_job(x->cores[0], 1, n, Abuf[q], 1, n, int(n), int(R1)));
      else
        gb.add(int(rp), int(n * R1), int(R0), 1.0, FR[q], 1, rp, x->cores[j], 1, R0, 0.0, Abuf[q], 1, rp);
      qj.push_back(QRJob{Abuf[q], rp * n, int(rp * n), int(R1), outs[q]->cores[j], rp * n});
    }
    copy2d_batched(ctx, cp);
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "orth_lr");
    qr_batched(ctx, qj);
    cp.clear();
    for (size_t q = 0; q < nb; ++q) {
      const int64_t n = xs[q]->dims[j], rp = yr[q][j], R1 = xs[q]->rks[j + 1], rn = yr[q][j + 1];
      cp.push_back(copy_job(Abuf[q], 1, rp * n, FR[q], 1, rn, int(rn), int(R1), 1.0, 1));
    }
    copy2d_batched(ctx, cp);
  }
  // ---- right-to-left (cores d-1..root): T = (X_j FL)_R^T, QR(T) = Q1 R1, core = Q1^T, FL = R1^T
  for (int j = maxd - 1; j >= root; --j) {
    std::vector<QRJob> qj;
    for (size_t q = 0; q < nb; ++q) {
      const DevTT* x = xs[q];
      if (j >= x->d) continue;
      const int64_t n = x->dims[j], R0 = x->rks[j], R1 = x->rks[j + 1], rn = yr[q][j + 1];
      const double* fl = (j == x->d - 1) ? h->one : FL[q];
      const int64_t ldfl = (j == x->d - 1) ? 1 : R1;
      const int64_t ldT = n * rn;
      // rows of T in the reference's order (src/tt_rounding.jl:85-86): T(b + rn i, a) = sum_g FL(g, b) X(a, i, g)
      for (int64_t i = 0; i < n; ++i)
        gb.add(int(rn), int(R0), int(R1), 1.0, fl, ldfl, 1, x->cores[j] + R0 * i, R0 * n, 1, 0.0, Abuf[q] + rn * i, 1,
               ldT);
      qj.push_back(QRJob{Abuf[q], ldT, int(ldT), int(R0), Qw[q], ldT});
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "orth_rl");
    qr_batched(ctx, qj);
    std::vector<Copy2D> cp;
    for (size_t q = 0; q < nb; ++q) {
      const DevTT* x = xs[q];
      if (j >= x->d) continue;
      const int64_t n = x->dims[j], R0 = x->rks[j], rn = yr[q][j + 1], k = yr[q][j];
      const int64_t ldT = n * rn;
      for (int64_t i = 0; i < n; ++i)  // core(c, i, b) = Q1(b + rn i, c)
        cp.push_back(copy_job(Qw[q] + rn * i, 1, ldT, outs[q]->cores[j] + k * i, k * n, 1, int(rn), int(k)));
      cp.push_back(copy_job(Abuf[q], 1, ldT, FL[q], R0, 1, int(k), int(R0), 1.0, 1));          // FL = R1^T
    }
    copy2d_batched(ctx, cp);
  }
  // ---- root core: FR X_root FL
  for (size_t q = 0; q < nb; ++q) {
    const DevTT* x = xs[q];
    const int i0 = root - 1;
    const int64_t n = x->dims[i0], R0 = x->rks[i0], R1 = x->rks[i0 + 1], r0 = yr[q][i0], r1 = yr[q][i0 + 1];
    const double* fr = (i0 == 0) ? h->one : FR[q];
    const double* fl = (i0 == x->d - 1) ? h->one : FL[q];
    gb.add(int(r0), int(n * R1), int(R0), 1.0, fr, 1, (i0 == 0) ? 1 : r0, x->cores[i0], 1, R0, 0.0, Abuf[q], 1, r0);
    (void)fl;
  }
  gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "orth_root");
  for (size_t q = 0; q < nb; ++q) {
    const DevTT* x = xs[q];
    const int i0 = root - 1;
    const int64_t n = x->dims[i0], R1 = x->rks[i0 + 1], r0 = yr[q][i0], r1 = yr[q][i0 + 1];
    const double* fl = (i0 == x->d - 1) ? h->one : FL[q];
    const int64_t ldfl = (i0 == x->d - 1) ? 1 : R1;
    gb.add(int(r0 * n), int(r1), int(R1), 1.0, Abuf[q], 1, r0 * n, fl, 1, ldfl, 0.0, outs[q]->cores[i0], 1, r0 * n);
  }
  gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "orth_root");
  return outs;
}

// ---------------------------------------------------------- SVD rounding sweep
struct SweepState {
  DevTT* y;                         // orthogonalised input (not modified)
  std::unique_ptr<DevTT> out;       // result, cores filled as the sweep proceeds
  std::vector<double*> cur;         // current core pointers
  std::vector<int64_t> rk;          // current ranks
  double* ping[2];
  int pp = 0;
  double *Mt, *Qb, *B, *J, *sv, *US;
  double *Jh, *USh;  // re-orthonormalised kept rotations and the matching (B J) columns
  int32_t* idx;
  int32_t* info;
  bool zero = false;
};

// _tt_rounding (src/tt_rounding.jl:174-269) on a batch of orthogonalised TTs
void svd_sweep(ttr_handle_s* h, std::vector<SweepState>& st, double tol, int64_t rmax, int direction) {
  Ctx& ctx = h->ctx;
  const size_t nb = st.size();
  int maxd = 0;
  for (auto& s : st) maxd = std::max(maxd, s.y->d);
  // zero guard on the root core
  {
    std::vector<const double*> ptrs;
    std::vector<int64_t> cnt;
    for (auto& s : st) {
      const int k = direction == TTR_DIR_LEFT ? s.y->d - 1 : 0;
      ptrs.push_back(s.y->cores[k]);
      cnt.push_back(s.y->core_elems(k));
    }
    std::vector<double> nr = frob_norms_host(h, ptrs, cnt);
    for (size_t q = 0; q < nb; ++q) st[q].zero = nr[q] < tol;
  }
  // singular values of every item in one array (one download per step)
  std::vector<int64_t> sv_off(nb + 1, 0);
  for (size_t q = 0; q < nb; ++q) {
    const DevTT* y = st[q].y;
    int64_t mxp = 1;
    for (int k = 0; k < y->d; ++k) {
      mxp = std::max(mxp, std::min(y->rks[k], y->dims[k] * y->rks[k + 1]));
      mxp = std::max(mxp, std::min(y->rks[k] * y->dims[k], y->rks[k + 1]));
    }
    sv_off[q + 1] = sv_off[q] + mxp;
  }
  double* sv_all = ctx.ws.alloc_n<double>(size_t(sv_off[nb]));
  std::vector<double> sv_host(static_cast<size_t>(sv_off[nb]));
  for (size_t q = 0; q < nb; ++q) st[q].sv = sv_all + sv_off[q];
  int64_t mxp_all = 1;
  for (auto& s : st) {
    s.cur = s.y->cores;
    s.rk = s.y->rks;
    int64_t mx = 1, mxp = 1;
    for (int k = 0; k < s.y->d; ++k) {
      mx = std::max(mx, s.y->core_elems(k));
      mxp = std::max(mxp, std::min(s.y->rks[k], s.y->dims[k] * s.y->rks[k + 1]));
      mxp = std::max(mxp, std::min(s.y->rks[k] * s.y->dims[k], s.y->rks[k + 1]));
    }
    s.ping[0] = ctx.ws.alloc_n<double>(size_t(mx));
    s.ping[1] = ctx.ws.alloc_n<double>(size_t(mx));
    s.Mt = ctx.ws.alloc_n<double>(size_t(mx));
    s.Qb = ctx.ws.alloc_n<double>(size_t(mx));
    s.US = ctx.ws.alloc_n<double>(size_t(mx));
    s.B = ctx.ws.alloc_n<double>(size_t(mxp * mxp));
    s.J = ctx.ws.alloc_n<double>(size_t(mxp * mxp));
    s.Jh = ctx.ws.alloc_n<double>(size_t(mxp * mxp));
    s.USh = ctx.ws.alloc_n<double>(size_t(mxp * mxp));
    mxp_all = std::max(mxp_all, mxp);
    s.idx = ctx.ws.alloc_n<int32_t>(size_t(mxp));
    s.out = h->make_tt_lazy(s.y->dims);
  }
  // an identity matrix (ld mxp_all) whose leading k x k block serves every item
  double* eyeK = ctx.ws.alloc_n<double>(size_t(mxp_all * mxp_all));
  fill2d_batched(ctx, {Fill2D{eyeK, mxp_all, int(mxp_all), int(mxp_all), 1.0}});
  // Jacobi counters of all items in one array (zeroed before every SVD step)
  int32_t* info_all = ctx.ws.alloc_n<int32_t>(size_t(8) * nb);
  // certificate flags of all items (device and host copies)
  int32_t* cflags = ctx.ws.alloc_n<int32_t>(nb);
  int32_t* hcflags = static_cast<int32_t*>(ctx.pin.alloc(sizeof(int32_t) * nb));
  for (size_t q = 0; q < nb; ++q) st[q].info = info_all + 8 * q;
  GemmBatch gb;
  const bool left = direction == TTR_DIR_LEFT;
  const int nsteps = maxd - 1;
  for (int step = 0; step < nsteps; ++step) {
    // core index being decomposed and its neighbour
    std::vector<int> jcore(nb, -1);
    std::vector<int64_t> P(nb), Qd(nb);
    std::vector<char> tallcase(nb, 0);
    std::vector<QRJob> qj;
    std::vector<Copy2D> cp;
    for (size_t q = 0; q < nb; ++q) {
      SweepState& s = st[q];
      const int d = s.y->d;
      if (s.zero || step >= d - 1) continue;
      const int j = left ? d - 1 - step : step;
      jcore[q] = j;
      const int64_t n = s.y->dims[j];
      int64_t p, qq;
      if (left) { p = s.rk[j]; qq = n * s.rk[j + 1]; }   // right unfolding p x q
      else { p = s.rk[j] * n; qq = s.rk[j + 1]; }        // left unfolding p x q
      P[q] = p;
      Qd[q] = qq;
      // left: wide (p <= q) -> QR of M^T;  right: tall (p >= q) -> QR of M
      const bool qr_of_m = left ? (p > qq) : (p >= qq);
      tallcase[q] = qr_of_m;
      if (qr_of_m) {
        cp.push_back(copy_job(s.cur[j], 1, p, s.Mt, 1, p, int(p), int(qq)));
        qj.push_back(QRJob{s.Mt, p, int(p), int(qq), s.Qb, p});
      } else {
        cp.push_back(copy_job(s.cur[j], 1, p, s.Mt, qq, 1, int(p), int(qq)));
        qj.push_back(QRJob{s.Mt, qq, int(qq), int(p), s.Qb, qq});
      }
    }
    if (qj.empty()) break;
    copy2d_batched(ctx, cp);
    qr_batched(ctx, qj);
    // B from the triangular factor: R (q x q) or R^T, see the per-case algebra below
    cp.clear();
    std::vector<int64_t> pdim(nb, 0);
    for (size_t q = 0; q < nb; ++q) {
      if (jcore[q] < 0) continue;
      SweepState& s = st[q];
      const int64_t p = P[q], qq = Qd[q];
      const bool t = tallcase[q];
      const int64_t m = t ? qq : p;  // size of the square factor
      const int64_t ldm = t ? p : qq;
      pdim[q] = m;
      // left&tall: B = R; left&wide: B = R^T; right&tall: B = R^T; right&wide: B = R
      const bool transpose = left ? !t : t;
      if (transpose) cp.push_back(copy_job(s.Mt, 1, ldm, s.B, m, 1, int(m), int(m), 1.0, 1));
      else cp.push_back(copy_job(s.Mt, 1, ldm, s.B, 1, m, int(m), int(m), 1.0, 1));
    }
    copy2d_batched(ctx, cp);
    // Bonds whose floor rule provably keeps every singular value (Cholesky
    // certificate on B^T B, written into J's buffer) need no SVD: the QR/LQ factor
    // already is an orthonormal basis of the kept space, and B is absorbed as it is.
    // Only bonds with m <= rmax qualify (rmax truncates the others).
    std::vector<char> cert(nb, 0);
    if (ctx.svd_certify) {
      std::vector<CholJob> cj;
      for (size_t q = 0; q < nb; ++q) {
        if (jcore[q] < 0) continue;
        const int64_t m = pdim[q];
        if (m > rmax || m > CHOL_MAX_M) continue;
        SweepState& s = st[q];
        gb.add(int(m), int(m), int(m), 1.0, s.B, m, 1, s.B, 1, m, 0.0, s.J, 1, m);
        cj.push_back(CholJob{s.J, m, int(m), tol, cflags + q});
      }
      if (!cj.empty()) {
        gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "svd_cert_gram");
        zero_fill(ctx, cflags, sizeof(int32_t) * nb);
        chol_certify_batched(ctx, cj);
        TTR_HIP(hipMemcpyAsync(hcflags, cflags, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, ctx.stream));
        ctx.sync();
        int64_t ncert = 0;
        for (size_t q = 0; q < nb; ++q) {
          cert[q] = jcore[q] >= 0 && hcflags[q] == 1;
          ncert += cert[q];
        }
        if (ctx.prof.on) {
          ProfStat& ps = ctx.prof.stats["svd_certify:bonds"];
          ps.launches += int64_t(cj.size());
          ps.flops += double(ncert);
        }
      }
    }
    std::vector<SVDJob> sj;
    for (size_t q = 0; q < nb; ++q) {
      if (jcore[q] < 0 || cert[q]) continue;
      SweepState& s = st[q];
      sj.push_back(SVDJob{s.B, pdim[q], s.J, pdim[q], s.sv, int(pdim[q]), s.info});
    }
    double wc = 64.0;
    if (!sj.empty()) {
      zero_fill(ctx, info_all, sizeof(int32_t) * 8 * nb);
      wc = jacobi_svd_batched(ctx, sj);
    }
    // rank decisions on the host
    ctx.sync();
    if (!sj.empty()) TTR_HIP(hipMemcpy(sv_host.data(), sv_all, sv_host.size() * 8, hipMemcpyDeviceToHost));
    std::vector<uint32_t> info_host;
    if (ctx.prof.on && !sj.empty()) {
      info_host.resize(8 * nb);
      TTR_HIP(hipMemcpy(info_host.data(), info_all, info_host.size() * 4, hipMemcpyDeviceToHost));
    }
    std::vector<int64_t> K(nb, 0);
    std::vector<int32_t> orders;  // the kept columns of every item, one upload
    std::vector<size_t> order_off(nb, 0);
    for (size_t q = 0; q < nb; ++q) {
      if (jcore[q] < 0) continue;
      const int64_t m = pdim[q];
      if (cert[q]) {  // every column kept, in place: B itself is (B J)_k, J = I
        K[q] = m;
        continue;
      }
      const double* sv = sv_host.data() + sv_off[q];
      if (ctx.prof.on && ctx.prof.level == 1) {
        const uint32_t* info = info_host.data() + 8 * q;
        ProfStat& ps = ctx.prof.stats["jacobi_svd:sweeps"];
        ps.launches += 1;
        ps.flops += info[0];
        ps.bytes = std::max(ps.bytes, double(info[0]));
        // visits and wall-clock ticks (100 MHz) summed over visits: Gram, inner sweep, apply
        ctx.prof.stats["jacobi_svd:visits"].flops += info[4];
        // MFMA work: every visit forms the upper triangle of a wc-column Gram (its
        // wc/16 (wc/16 + 1) / 2 tiles of 16 x 16, 2*256*m flops each), a rotating
        // visit applies V to wc columns of B and of J (2 * 2*m*wc*wc); bytes: the
        // columns read by the Gram, read and written by the apply
        const double gram_tiles = (wc / 16) * (wc / 16 + 1) / 2;
        ctx.prof.stats["jacobi_svd"].flops +=
            2.0 * m * (256.0 * gram_tiles * double(info[4]) + 2.0 * wc * wc * info[5]);
        ctx.prof.stats["jacobi_svd"].bytes += 8.0 * wc * m * (double(info[4]) + 4.0 * info[5]);
        // algorithmic work of the same SVD: 21 m^3, the Golub-Kahan-Reinsch count for
        // Sigma, U and V of a square matrix (Golub and Van Loan, Matrix Computations)
        ctx.prof.stats["jacobi_svd:alg_flops"].flops += 21.0 * double(m) * m * m;
        ctx.prof.stats["jacobi_svd:ticks_gram"].flops += info[1];
        ctx.prof.stats["jacobi_svd:ticks_inner"].flops += info[2];
        ctx.prof.stats["jacobi_svd:ticks_apply"].flops += info[3];
      }
      std::vector<int32_t> order(m);
      std::iota(order.begin(), order.end(), 0);
      std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return sv[a] > sv[b]; });
      std::vector<double> ss(m);
      int64_t npos = 0;
      for (int64_t i = 0; i < m; ++i) {
        ss[i] = sv[order[i]];
        npos += ss[i] > 0.0;
      }
      if (!std::all_of(ss.begin(), ss.end(), [](double v) { return std::isfinite(v); }))
        throw Error(ST_NUMERICAL_FAILURE, "SVD produced non-finite singular values");
      int64_t k = floor_rule(ss, tol);
      k = std::min<int64_t>({k, rmax, npos});
      k = std::max<int64_t>(k, 1);
      K[q] = k;
      if (ctx.prof.on) {  // SVD bonds, and those whose rank decision dropped something
        ProfStat& ps = ctx.prof.stats["svd_sweep:svd_bonds"];
        ps.launches += 1;
        ps.flops += k < m ? 1.0 : 0.0;
      }
      order_off[q] = orders.size();
      orders.insert(orders.end(), order.begin(), order.begin() + k);
    }
    const int32_t* dorders = h->ctx.upload(orders);
    // reorder columns of B*J and J in place order (into US / Mt scratch); the
    // second gather must not overwrite B before the first reads it: two launches
    std::vector<GatherCols> g1, g2;
    for (size_t q = 0; q < nb; ++q) {
      if (jcore[q] < 0 || cert[q]) continue;
      SweepState& s = st[q];
      const int64_t m = pdim[q], k = K[q];
      int32_t* didx = const_cast<int32_t*>(dorders) + order_off[q];
      g1.push_back(GatherCols{s.B, s.US, m, m, didx, int(m), int(k)});
      g2.push_back(GatherCols{s.J, s.B, m, m, didx, int(m), int(k)});
    }
    gather_cols_batched(ctx, g1);  // US <- (B J)[:, sel]        (m x k)
    // J[:, sel] goes to a fresh buffer (B is free now)
    gather_cols_batched(ctx, g2);  // B  <- J[:, sel]            (m x k)
    // The accumulated Jacobi rotations drift from orthogonality by about 64 eps per
    // applied visit (tens of visits per column); across a sweep of truncating bonds
    // that drift, not the SVD's accuracy, dominates the error.  One Newton-Schulz
    // step C = (3 I - J_k^T J_k) / 2 makes Jh = J_k C orthonormal to rounding (the
    // deviation is squared), and the same C on (B J)_k keeps the represented
    // product (B J_k C)(J_k C)^T = B Jh Jh^T an orthogonal projection of B.
    for (size_t q = 0; q < nb; ++q) {
      if (jcore[q] < 0 || cert[q]) continue;
      SweepState& s = st[q];
      const int64_t m = pdim[q], k = K[q];
      gb.add(int(k), int(k), int(m), 1.0, s.B, m, 1, s.B, 1, m, 0.0, s.J, 1, k);  // G = J_k^T J_k
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "svd_orth_gram");
    for (size_t q = 0; q < nb; ++q) {
      if (jcore[q] < 0 || cert[q]) continue;
      SweepState& s = st[q];
      const int64_t m = pdim[q], k = K[q];
      for (int w = 0; w < 2; ++w) {  // Jh = J_k C, USh = (B J)_k C
        const double* src = w == 0 ? s.B : s.US;
        gb.prob(w == 0 ? s.Jh : s.USh, 1, m, int(m), int(k), 1.0, 0.0);
        gb.seg(src, 1, m, eyeK, 1, mxp_all, int(k), 1.5);
        gb.seg(src, 1, m, s.J, 1, k, int(k), -0.5);
      }
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "svd_orth_fix");
    // build the new cores (a certified bond: (B J)_k = B, J_k = I)
    std::vector<Copy2D> cpo;
    std::vector<Fill2D> eye;
    for (size_t q = 0; q < nb; ++q) {
      if (jcore[q] < 0) continue;
      SweepState& s = st[q];
      const int j = jcore[q];
      const int64_t p = P[q], qq = Qd[q], k = K[q], m = pdim[q];
      const bool t = tallcase[q];
      // (m x k): columns u_c s_c of the square factor, and the right rotations
      const double* BJk = cert[q] ? s.B : s.USh;
      const double* Jk = s.Jh;
      if (left) {
        const int64_t n = s.y->dims[j];
        double* newj = h->core_alloc(s.out.get(), k * qq);
        const int jl = j - 1;
        const int64_t rl = s.rk[jl] * s.y->dims[jl];
        double* dst = s.ping[s.pp];
        if (!t) {
          // M = (B J)(Q J)^T: core_j = (Q Jk)^T, core_{j-1} <- core_{j-1} (B J)_k
          if (cert[q]) cpo.push_back(copy_job(s.Qb, qq, 1, newj, 1, k, int(k), int(qq)));  // J = I
          else gb.add(int(k), int(qq), int(m), 1.0, Jk, m, 1, s.Qb, qq, 1, 0.0, newj, 1, k);
          gb.add(int(rl), int(k), int(p), 1.0, s.cur[jl], 1, rl, BJk, 1, m, 0.0, dst, 1, rl);
        } else {
          // M = Q (R J) J^T: core_j = Jk^T, core_{j-1} <- core_{j-1} (Q (R J)_k)
          if (cert[q]) eye.push_back(Fill2D{newj, k, int(k), int(m), 1.0});
          else cpo.push_back(copy_job(Jk, 1, m, newj, k, 1, int(m), int(k)));
          gb.add(int(p), int(k), int(qq), 1.0, s.Qb, 1, p, BJk, 1, m, 0.0, s.Mt, 1, p);
        }
        (void)n;
        s.cur[j] = newj;
        s.out->cores[j] = newj;
      } else {
        const int jr = j + 1;
        const int64_t nr = s.y->dims[jr] * s.rk[jr + 1];
        double* newj = h->core_alloc(s.out.get(), p * k);
        double* dst = s.ping[s.pp];
        if (t) {
          // M = Q J (R^T J)^T... : U = Q Jk, S V^T = (B J)_k^T
          if (cert[q]) cpo.push_back(copy_job(s.Qb, 1, p, newj, 1, p, int(p), int(k)));  // J = I
          else gb.add(int(p), int(k), int(qq), 1.0, s.Qb, 1, p, Jk, 1, m, 0.0, newj, 1, p);
          gb.add(int(k), int(nr), int(qq), 1.0, BJk, m, 1, s.cur[jr], 1, qq, 0.0, dst, 1, k);
        } else {
          // M = J (Q (R J))^T: U = Jk, S V^T = (Q (B J)_k)^T
          if (cert[q]) eye.push_back(Fill2D{newj, p, int(p), int(k), 1.0});
          else cpo.push_back(copy_job(Jk, 1, m, newj, 1, p, int(p), int(k)));
          gb.add(int(k), int(qq), int(p), 1.0, BJk, m, 1, s.Qb, qq, 1, 0.0, s.Mt, 1, k);
        }
        s.cur[j] = newj;
        s.out->cores[j] = newj;
      }
    }
    copy2d_batched(ctx, cpo);
    fill2d_batched(ctx, eye);
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "svd_sweep_a");
    // second-stage products of the tall-left / wide-right cases
    for (size_t q = 0; q < nb; ++q) {
      if (jcore[q] < 0) continue;
      SweepState& s = st[q];
      const int j = jcore[q];
      const int64_t p = P[q], qq = Qd[q], k = K[q];
      const bool t = tallcase[q];
      double* dst = s.ping[s.pp];
      if (left && t) {
        const int jl = j - 1;
        const int64_t rl = s.rk[jl] * s.y->dims[jl];
        gb.add(int(rl), int(k), int(p), 1.0, s.cur[jl], 1, rl, s.Mt, 1, p, 0.0, dst, 1, rl);
      } else if (!left && !t) {
        const int jr = j + 1;
        const int64_t nr = s.y->dims[jr] * s.rk[jr + 1];
        gb.add(int(k), int(nr), int(qq), 1.0, s.Mt, 1, k, s.cur[jr], 1, qq, 0.0, dst, 1, k);
      }
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "svd_sweep_b");
    for (size_t q = 0; q < nb; ++q) {
      if (jcore[q] < 0) continue;
      SweepState& s = st[q];
      const int j = jcore[q];
      if (left) {
        s.cur[j - 1] = s.ping[s.pp];
        s.rk[j] = K[q];
        s.out->ot[j] = -1;
      } else {
        s.cur[j + 1] = s.ping[s.pp];
        s.rk[j + 1] = K[q];
        s.out->ot[j] = 1;
      }
      s.pp ^= 1;
    }
  }
  // finish: copy the last working core, set metadata
  std::vector<Copy2D> cp;
  for (auto& s : st) {
    const int d = s.y->d;
    if (s.zero) {
      s.out = h->make_tt(s.y->dims, s.y->rks, std::vector<int64_t>(d, 0));
      zeros_like_ranks(h, s.out.get());
      continue;
    }
    const int jf = left ? 0 : d - 1;
    s.out->rks = s.rk;
    const int64_t e = s.out->core_elems(jf);
    double* c = h->core_alloc(s.out.get(), e);
    cp.push_back(copy_job(s.cur[jf], 1, e, c, 1, e, int(e), 1));
    s.out->cores[jf] = c;
    if (left) s.out->ot[0] = 0;
    else s.out->ot[d - 1] = 0;
  }
  copy2d_batched(ctx, cp);
}

// ------------------------------------------------------------ randomized
struct SketchPlan {
  int d = 0;
  std::vector<int64_t> dims, bv, pmax;
  Prec prec = Prec::F64;   // the blocks
  Prec wprec = Prec::F64;  // the sketch tensors W_k, their temporaries and the contractions' math
  bool orthogonal = true;
  uint64_t seed = 1234;
  const double* const* blocks = nullptr;  // caller-supplied blocks (reference layout), or null
  std::vector<void*> S;  // per core, device layout (n, bvr, bvl, pmax)
};

std::vector<int64_t> block_counts(const std::vector<int64_t>& rks, const std::vector<int64_t>& bv, int d,
                                  const ttr_rand_opts* o) {
  std::vector<int64_t> p(d + 1, 1);
  if (o->p) {
    for (int k = 0; k <= d; ++k) p[k] = o->p[k];
  } else {
    // compute_sketch_blocks_heuristic(...; reverse=true) (src/sketches.jl:107-123)
    for (int k = 0; k <= d; ++k) p[k] = int64_t(std::ceil(double(rks[k]) / double(bv[k])));
    for (int k = 1; k < d; ++k) p[k] = std::max(p[k], p[k - 1]);
    p[d] = 1;
  }
  if (o->oversampling != 1.0 && o->oversampling > 0.0)
    for (int k = 0; k < d; ++k) p[k] = int64_t(std::ceil(o->oversampling * double(p[k])));
  for (int k = 0; k <= d; ++k) TTR_REQUIRE(p[k] >= 1, ST_INVALID_ARG, "block counts must be >= 1");
  return p;
}

std::vector<int64_t> block_ranks(const std::vector<int64_t>& dims, int64_t block_rks) {
  const int d = int(dims.size());
  std::vector<int64_t> bv(d + 1, 1);
  for (int k = 0; k < d; ++k) bv[k] = block_rks;
  for (int k = d - 1; k >= 0; --k) bv[k] = std::min(bv[k], dims[k] * bv[k + 1]);
  return bv;
}

// generate_sketch_blocks (src/sketches.jl:44-84) for cores kmin..d-1
void make_sketch(ttr_handle_s* h, SketchPlan& sp, int kmin) {
  Ctx& ctx = h->ctx;
  const size_t eb = prec_bytes(sp.prec);
  sp.S.assign(sp.d, nullptr);
  if (sp.blocks) {  // the caller's blocks: upload and pack into the device layout
    for (int k = kmin; k < sp.d; ++k) {
      TTR_REQUIRE(sp.blocks[k] != nullptr, ST_INVALID_ARG, "blocks[" + std::to_string(k) + "] is NULL");
      const int64_t n = sp.dims[k], bvr = sp.bv[k + 1], bvl = sp.bv[k], p = sp.pmax[k];
      const int64_t cnt = n * bvr * bvl * p;
      sp.S[k] = ctx.ws.alloc(size_t(cnt) * eb);
      const double* dref = ctx.upload(sp.blocks[k], size_t(cnt));
      sketch_pack(ctx, dref, sp.S[k], int(n), int(bvr), int(bvl), int(p), 1.0, sp.prec);
    }
    return;
  }
  std::vector<SketchGen> gen;
  for (int k = kmin; k < sp.d; ++k) {
    const int64_t n = sp.dims[k], bvr = sp.bv[k + 1], bvl = sp.bv[k], p = sp.pmax[k];
    const int64_t cnt = n * bvr * bvl * p;
    sp.S[k] = ctx.ws.alloc(size_t(cnt) * eb);
    if (!sp.orthogonal) {
      gen.push_back(SketchGen{sp.S[k], nullptr, int(n), int(bvr), int(bvl), int(p), k, 1.0 / std::sqrt(double(bvl))});
    }
  }
  if (!sp.orthogonal) {
    sketch_generate(ctx, gen, sp.seed, sp.prec);
    return;
  }
  // orthogonal: FP64 blocks in reference order, thin Householder Q per block (or identity), scale, pack
  std::vector<double*> ref(sp.d, nullptr), qb(sp.d, nullptr);
  std::vector<Fill2D> fills;
  std::vector<QRJob> qj;
  for (int k = kmin; k < sp.d; ++k) {
    const int64_t n = sp.dims[k], bvr = sp.bv[k + 1], bvl = sp.bv[k], p = sp.pmax[k];
    const int64_t m = bvr * n;
    ref[k] = ctx.ws.alloc_n<double>(size_t(m * bvl * p));
    if (bvl == m) {
      for (int64_t j = 0; j < p; ++j) fills.push_back(Fill2D{ref[k] + j * m * bvl, m, int(m), int(bvl), 1.0});
    } else {
      gen.push_back(SketchGen{nullptr, ref[k], int(n), int(bvr), int(bvl), int(p), k, 1.0});
      qb[k] = ctx.ws.alloc_n<double>(size_t(m * bvl * p));
      for (int64_t j = 0; j < p; ++j)
        qj.push_back(QRJob{ref[k] + j * m * bvl, m, int(m), int(bvl), qb[k] + j * m * bvl, m});
    }
  }
  fill2d_batched(ctx, fills);
  sketch_generate(ctx, gen, sp.seed, Prec::F64);
  qr_batched(ctx, qj);
  for (int k = kmin; k < sp.d; ++k) {
    const int64_t n = sp.dims[k], bvr = sp.bv[k + 1], bvl = sp.bv[k], p = sp.pmax[k];
    const int64_t m = bvr * n;
    if (bvl == m) sketch_pack(ctx, ref[k], sp.S[k], int(n), int(bvr), int(bvl), int(p), 1.0, sp.prec);
    else sketch_pack(ctx, qb[k], sp.S[k], int(n), int(bvr), int(bvl), int(p),
                     std::sqrt(double(m) / double(bvl)), sp.prec);
  }
}

// One term alpha * x, or alpha * H x when H is set (implicit operator product:
// its bond k is the pair (x rank, H rank), x rank fastest, of size rc(k)).
struct RandTerm {
  DevTT* x;
  double alpha;
  DevTTO* H = nullptr;
  std::vector<void*> W;     // per core k: (rc(k) x s_k) in sketch precision
  int32_t* E;               // per core exponent (d+1 ints, device)
  double* U = nullptr;      // operator terms: P x_{k+1} for every H bond
  int64_t rc(int k) const { return x->rks[k] * (H ? H->rks[k] : 1); }
};

// strides of H(out mode, out bond, in mode, in bond) in a reference-layout
// operator core (n, n, R, R'): towards the left (sketch: the right bond is
// contracted) or towards the right (projection: the left bond is contracted)
OpApply op_job(const DevTTO* H, int k, bool left_out) {
  OpApply j{};
  const int64_t n = H->dims[k], Rl = H->rks[k], Rr = H->rks[k + 1];
  j.H = H->cores[k];
  j.n = int(n);
  j.h_om = 1;
  j.h_im = n;
  if (left_out) {
    j.h_ob = n * n;
    j.h_ib = n * n * Rl;
    j.rout = int(Rl);
    j.rin = int(Rr);
  } else {
    j.h_ob = n * n * Rl;
    j.h_ib = n * n;
    j.rout = int(Rr);
    j.rin = int(Rl);
  }
  j.scale = 1.0;
  return j;
}

struct RandItem {
  std::vector<RandTerm> terms;
  std::vector<int64_t> rks, p, s;  // target ranks, block counts, sketch ranks (d+1)
  double* Z = nullptr;
  // The terms' current cores side by side, [Y_1 | Y_2 | ...] ((r_k n_k) x sum_t rc_t(k+1),
  // two buffers for the current and the next core), and their projections
  // [P_1 | P_2 | ...] (r_{k+1} x sum_t rc_t(k+1)): one GEMM forms every P_t of an item
  double* Yall[2] = {nullptr, nullptr};
  double* Pall = nullptr;
  // column of term t's block at bond k
  int64_t col(size_t t, int k) const {
    int64_t c = 0;
    for (size_t u = 0; u < t; ++u) c += terms[u].rc(k);
    return c;
  }
  int64_t cols(int k) const { return col(terms.size(), k); }
};

// right-to-left sketch phase of tt_recursive_sketch (src/sketches.jl:221-240);
// operator terms follow the implicit operator sketch (src/sketches.jl:392-455,
// contract_sketch_core_backwards! for operators :1648-1713): per core, T1 =
// X_k W_{k+1} (GEMM), T2 = the operator core applied to the (mode, right H bond)
// pairs of T1 (op_apply), W_k = T2 S_k (GEMM per block)
void sketch_phase(ttr_handle_s* h, SketchPlan& sp, std::vector<RandItem>& items, int kmin) {
  Ctx& ctx = h->ctx;
  const int d = sp.d;
  const size_t eb = prec_bytes(sp.wprec), es = prec_bytes(sp.prec);  // W / temporaries, blocks
  const Prec wp = sp.wprec;
  // buffers
  std::vector<std::vector<void*>> T1(items.size()), T2(items.size());
  // every term's sketch exponents in one array, cleared by one launch
  size_t nterm = 0;
  for (auto& it : items) nterm += it.terms.size();
  int32_t* E_all = ctx.ws.alloc_n<int32_t>(nterm * size_t(d + 1));
  zero_fill(ctx, E_all, nterm * size_t(d + 1) * 4);
  size_t term_idx = 0;
  for (size_t i = 0; i < items.size(); ++i) {
    for (auto& t : items[i].terms) {
      t.W.assign(d + 1, nullptr);
      t.E = E_all + (term_idx++) * size_t(d + 1);
      int64_t mx = 1, mx2 = 1;
      for (int k = kmin; k < d; ++k) {
        t.W[k] = ctx.ws.alloc(size_t(t.rc(k) * items[i].s[k]) * eb);
        const int64_t cols = sp.bv[k + 1] * items[i].p[k];
        if (k < d - 1) mx = std::max(mx, t.x->rks[k] * sp.dims[k] * (t.H ? t.H->rks[k + 1] : 1) * cols);
        if (t.H) mx2 = std::max(mx2, t.rc(k) * sp.dims[k] * cols);
      }
      T1[i].push_back(ctx.ws.alloc(size_t(mx) * eb));
      T2[i].push_back(t.H ? ctx.ws.alloc(size_t(mx2) * eb) : nullptr);
    }
  }
  GemmBatch gb;
  std::vector<OpApply> oj;
  for (int k = d - 1; k >= kmin; --k) {
    const int64_t n = sp.dims[k], bvr = sp.bv[k + 1], bvl = sp.bv[k];
    const char* S = static_cast<const char*>(sp.S[k]);
    if (k == d - 1) {
      for (auto& it : items)
        for (auto& t : it.terms) {
          if (t.H) continue;
          const int64_t Rk = t.x->rks[k];
          for (int64_t j = 0; j < it.p[k]; ++j)
            gb.add(int(Rk), int(bvl), int(n), 1.0, t.x->cores[k], 1, Rk, S + size_t(n * bvl * j) * es, 1, n, 0.0,
                   static_cast<char*>(t.W[k]) + size_t(Rk * bvl * j) * eb, 1, Rk);
        }
      gemm_run(ctx, gb, Prec::F64, sp.prec, wp, wp, "sketch_last");
      // operator terms: T2 (r R x n) = H_{d-1} applied to the last core, then T2 S_j
      for (size_t i = 0; i < items.size(); ++i)
        for (size_t ti = 0; ti < items[i].terms.size(); ++ti) {
          RandTerm& t = items[i].terms[ti];
          if (!t.H) continue;
          const int64_t r = t.x->rks[k], rR = t.rc(k);
          OpApply o = op_job(t.H, k, true);
          o.in = t.x->cores[k];
          o.out = T2[i][ti];
          o.X = int(r);
          o.Y = 1;
          o.is_m = r; o.is_b = 0; o.is_x = 1; o.is_y = 0;
          o.os_m = rR; o.os_b = r; o.os_x = 1; o.os_y = 0;
          oj.push_back(o);
          for (int64_t j = 0; j < items[i].p[k]; ++j)
            gb.add(int(rR), int(bvl), int(n), 1.0, T2[i][ti], 1, rR, S + size_t(n * bvl * j) * es, 1, n, 0.0,
                   static_cast<char*>(t.W[k]) + size_t(rR * bvl * j) * eb, 1, rR);
        }
      op_apply_batched(ctx, oj, Prec::F64, wp);
      oj.clear();
      gemm_run(ctx, gb, wp, sp.prec, wp, wp, "sketch_last_op");
    } else {
      for (size_t i = 0; i < items.size(); ++i)
        for (size_t ti = 0; ti < items[i].terms.size(); ++ti) {
          RandTerm& t = items[i].terms[ti];
          const int64_t Rk = t.x->rks[k], Rk1 = t.x->rks[k + 1], p = items[i].p[k];
          const int64_t Hr1 = t.H ? t.H->rks[k + 1] : 1;
          // W_{k+1} (Rk1*Hr1 x s) read as Rk1 x (Hr1 * s): columns (H bond, sketch column)
          gb.add(int(Rk * n), int(Hr1 * bvr * p), int(Rk1), 1.0, t.x->cores[k], 1, Rk * n, t.W[k + 1], 1, Rk1, 0.0,
                 T1[i][ti], 1, Rk * n);
          if (t.H) {
            // T2[(a, b), (zeta, c)] = sum_{z, beta} H[zeta, z, b, beta] T1[(a, z), (beta, c)]
            const int64_t rR = t.rc(k);
            OpApply o = op_job(t.H, k, true);
            o.in = T1[i][ti];
            o.out = T2[i][ti];
            o.X = int(Rk);
            o.Y = int(bvr * p);
            o.is_m = Rk; o.is_b = Rk * n; o.is_x = 1; o.is_y = Rk * n * Hr1;
            o.os_m = rR; o.os_b = Rk; o.os_x = 1; o.os_y = rR * n;
            oj.push_back(o);
          }
        }
      gemm_run(ctx, gb, Prec::F64, wp, wp, wp, "sketch_t1");
      op_apply_batched(ctx, oj, wp, wp);
      oj.clear();
      for (size_t i = 0; i < items.size(); ++i)
        for (size_t ti = 0; ti < items[i].terms.size(); ++ti) {
          RandTerm& t = items[i].terms[ti];
          const int64_t Rk = t.rc(k);
          const char* A = static_cast<const char*>(t.H ? T2[i][ti] : T1[i][ti]);
          for (int64_t j = 0; j < items[i].p[k]; ++j)
            gb.add(int(Rk), int(bvl), int(n * bvr), 1.0, A + size_t(Rk * n * bvr * j) * eb, 1, Rk,
                   S + size_t(n * bvr * bvl * j) * es, 1, n * bvr, 0.0,
                   static_cast<char*>(t.W[k]) + size_t(Rk * bvl * j) * eb, 1, Rk);
        }
      gemm_run(ctx, gb, wp, sp.prec, wp, wp, "sketch_w");
    }
    std::vector<Pow2Job> pj;
    for (auto& it : items)
      for (auto& t : it.terms)
        pj.push_back(Pow2Job{t.W[k], t.rc(k), int(t.rc(k)), int(it.s[k]), k == d - 1 ? nullptr : t.E + k + 1,
                             t.E + k});
    pow2_normalize_batched(ctx, pj, wp);
  }
}

void prepare_plan(ttr_handle_s* h, const ttr_rand_opts* o, int d, const std::vector<int64_t>& dims,
                  std::vector<RandItem>& items, SketchPlan& sp) {
  TTR_REQUIRE(o != nullptr, ST_INVALID_ARG, "options must not be NULL");
  TTR_REQUIRE(o->forward == 0, ST_INVALID_ARG, "this entry point takes the right-to-left sketch only");
  sp.d = d;
  sp.dims = dims;
  sp.prec = to_prec(o->sketch_prec);
  TTR_REQUIRE(o->sketch_accum == 0 || o->sketch_accum == 1, ST_INVALID_ARG, "sketch_accum must be 0 or 1");
  sp.wprec = o->sketch_accum ? Prec::F64 : sp.prec;
  sp.orthogonal = o->orthogonal != 0;
  sp.seed = o->seed;
  sp.blocks = o->blocks;
  sp.bv = block_ranks(dims, o->block_rks > 0 ? o->block_rks : d);
  sp.pmax.assign(d + 1, 1);
  for (auto& it : items) {
    TTR_REQUIRE(int(it.rks.size()) == d + 1 && it.rks[d] == 1, ST_INVALID_ARG, "rks must have d+1 entries with rks[d] == 1");
    it.p = block_counts(it.rks, sp.bv, d, o);
    it.s.resize(d + 1);
    for (int k = 0; k <= d; ++k) {
      it.s[k] = sp.bv[k] * it.p[k];
      sp.pmax[k] = std::max(sp.pmax[k], it.p[k]);
    }
    for (int k = 0; k + 1 < d; ++k)
      TTR_REQUIRE(it.p[k] <= it.p[k + 1], ST_INVALID_ARG, "block counts must be non-decreasing towards the right");
    TTR_REQUIRE(!sp.blocks || it.p == items[0].p, ST_INVALID_ARG,
                "with caller-supplied blocks every item must have the same block counts");
  }
  (void)h;
}

// ttrand_rounding(alpha, y, rks) (src/tt_randtools.jl:223-276), batched; a first
// term with an operator is alpha_1 H y_1 with the implicit operator sketch and
// projection of ttrand_rounding(alpha, A, y, rks) (:285-354)
std::vector<std::unique_ptr<DevTT>> round_rand_batch(ttr_handle_s* h, std::vector<RandItem>& items,
                                                     const ttr_rand_opts* o) {
  Ctx& ctx = h->ctx;
  const DevTT* x0 = items[0].terms[0].x;
  const int d = x0->d;
  for (auto& it : items)
    for (auto& t : it.terms) {
      TTR_REQUIRE(t.x->dims == x0->dims, ST_DIM_MISMATCH, "all terms must share dims");
      if (t.H) TTR_REQUIRE(t.H->dims == x0->dims, ST_DIM_MISMATCH, "operator and TT dims differ");
    }
  SketchPlan sp;
  prepare_plan(h, o, d, x0->dims, items, sp);
  for (auto& it : items)
    for (auto& t : it.terms) TTR_REQUIRE(t.x->rks[d] == 1, ST_INVALID_ARG, "rks[d] must be 1");
  std::vector<std::unique_ptr<DevTT>> outs(items.size());
  std::vector<std::vector<int64_t>> orr(items.size());
  for (size_t i = 0; i < items.size(); ++i) {
    std::vector<int64_t> r(d + 1, 1);
    for (int k = 0; k + 1 < d; ++k) r[k + 1] = std::min(sp.dims[k] * r[k], items[i].s[k + 1]);
    std::vector<int64_t> ot(d, 1);
    ot[d - 1] = 0;
    outs[i] = h->make_tt(sp.dims, r, ot);
    orr[i] = r;
  }
  GemmBatch gb;
  std::vector<OpApply> oj;
  if (d == 1) {
    // a single core: sum_t alpha_t (H_t) x_t
    for (size_t i = 0; i < items.size(); ++i)
      for (auto& t : items[i].terms) {
        if (!t.H) continue;
        OpApply a = op_job(t.H, 0, false);
        a.in = t.x->cores[0];
        a.out = outs[i]->cores[0];
        a.X = 1; a.Y = 1;
        a.is_m = 1; a.os_m = 1;
        a.scale = t.alpha;
        oj.push_back(a);
      }
    op_apply_batched(ctx, oj, Prec::F64, Prec::F64);
    for (size_t i = 0; i < items.size(); ++i) {
      bool op = false;
      for (auto& t : items[i].terms) op = op || t.H;
      bool any = false;
      for (auto& t : items[i].terms) {
        if (t.H) continue;
        if (!any) gb.prob(outs[i]->cores[0], 1, 1, 1, int(sp.dims[0]), 1.0, op ? 1.0 : 0.0);
        any = true;
        gb.seg(h->one, 1, 1, t.x->cores[0], 1, 1, 1, t.alpha);
      }
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "rr_d1");
    return outs;
  }
  {
    ProfScope ps(ctx, "phase:reverse_sketch");
    make_sketch(h, sp, 1);
    sketch_phase(h, sp, items, 1);
  }
  ProfScope ps_orth(ctx, "phase:orthogonalization");
  // buffers of the left-to-right sweep
  for (size_t i = 0; i < items.size(); ++i) {
    int64_t zmax = 1;
    for (int k = 0; k + 1 < d; ++k) zmax = std::max(zmax, orr[i][k] * sp.dims[k] * items[i].s[k + 1]);
    items[i].Z = ctx.ws.alloc_n<double>(size_t(zmax));
    int64_t ymax = 1, pmax = 1;
    for (int k = 0; k + 1 < d; ++k) {
      ymax = std::max(ymax, orr[i][k] * sp.dims[k] * items[i].cols(k + 1));
      pmax = std::max(pmax, orr[i][k + 1] * items[i].cols(k + 1));
    }
    items[i].Yall[0] = ctx.ws.alloc_n<double>(size_t(ymax));
    items[i].Yall[1] = ctx.ws.alloc_n<double>(size_t(ymax));
    items[i].Pall = ctx.ws.alloc_n<double>(size_t(pmax));
    for (auto& t : items[i].terms) {
      int64_t umax = 1;
      for (int k = 0; k + 1 < d; ++k)
        if (t.H) umax = std::max(umax, orr[i][k + 1] * sp.dims[k + 1] * t.x->rks[k + 2] * t.H->rks[k + 1]);
      if (t.H) t.U = ctx.ws.alloc_n<double>(size_t(umax));
    }
  }
  // term tt's core at step k in buffer b, and its projection P at step k
  auto Yat = [&](size_t i, size_t tt, int k, int b) {
    return items[i].Yall[b] + orr[i][k] * sp.dims[k] * items[i].col(tt, k + 1);
  };
  auto Pat = [&](size_t i, size_t tt, int k) { return items[i].Pall + orr[i][k + 1] * items[i].col(tt, k + 1); };
  // operator terms: the first core of H x (1 x n x (r_1 R_1), x rank fastest)
  for (size_t i = 0; i < items.size(); ++i)
    for (size_t tt = 0; tt < items[i].terms.size(); ++tt) {
      const RandTerm& t = items[i].terms[tt];
      if (!t.H) continue;
      const int64_t n0 = sp.dims[0], r1 = t.x->rks[1];
      OpApply a = op_job(t.H, 0, false);
      a.in = t.x->cores[0];
      a.out = Yat(i, tt, 0, 0);
      a.X = 1;
      a.Y = int(r1);
      a.is_m = 1; a.is_b = 0; a.is_x = 0; a.is_y = n0;
      a.os_m = 1; a.os_b = n0 * r1; a.os_x = 0; a.os_y = n0;
      oj.push_back(a);
    }
  op_apply_batched(ctx, oj, Prec::F64, Prec::F64);
  oj.clear();
  int yb = 0;
  for (int k = 0; k + 1 < d; ++k) {
    const int64_t n = sp.dims[k];
    auto Ycur = [&](size_t i, size_t tt) -> const double* {
      const RandTerm& t = items[i].terms[tt];
      return (k == 0 && !t.H) ? t.x->cores[0] : Yat(i, tt, k, yb);
    };
    // Z_k = sum_t Y_t W_t (FP64 accumulate; terms rescaled by their sketch exponents)
    for (size_t i = 0; i < items.size(); ++i) {
      const int64_t r = orr[i][k], s = items[i].s[k + 1];
      gb.prob(items[i].Z, 1, r * n, int(r * n), int(s), 1.0, 0.0);
      for (size_t tt = 0; tt < items[i].terms.size(); ++tt) {
        const RandTerm& t = items[i].terms[tt];
        const int64_t R1 = t.rc(k + 1);
        gb.seg(Ycur(i, tt), 1, r * n, t.W[k + 1], 1, R1, int(R1), k == 0 ? t.alpha : 1.0, AM_PLAIN, t.E + k + 1);
      }
    }
    gemm_run(ctx, gb, Prec::F64, sp.wprec, Prec::F64, Prec::F64, "rr_Z");
    std::vector<QRJob> qj;
    for (size_t i = 0; i < items.size(); ++i) {
      const int64_t r = orr[i][k], s = items[i].s[k + 1];
      qj.push_back(QRJob{items[i].Z, r * n, int(r * n), int(s), outs[i]->cores[k], r * n});
    }
    qr_batched(ctx, qj);
    // P_t = Q^T Y_t: per term on the first core (the terms' own cores, scaled by
    // alpha_t), then one product Q^T [Y_1 | Y_2 | ...] per item (same bits per entry)
    for (size_t i = 0; i < items.size(); ++i) {
      const int64_t r = orr[i][k], r1 = orr[i][k + 1];
      if (k > 0) {
        gb.add(int(r1), int(items[i].cols(k + 1)), int(r * n), 1.0, outs[i]->cores[k], r * n, 1, Yat(i, 0, k, yb), 1,
               r * n, 0.0, items[i].Pall, 1, r1);
        continue;
      }
      for (size_t tt = 0; tt < items[i].terms.size(); ++tt) {
        const RandTerm& t = items[i].terms[tt];
        const int64_t R1 = t.rc(k + 1);
        gb.prob(Pat(i, tt, k), 1, r1, int(r1), int(R1), 1.0, 0.0);
        gb.seg(outs[i]->cores[k], r * n, 1, Ycur(i, tt), 1, r * n, int(r * n), k == 0 ? t.alpha : 1.0);
      }
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "rr_P");
    // Y_t <- P_t X_{t,k+1}; the last core sums the terms.  Operator terms: U_beta =
    // P_beta X_{k+1} for every left H bond beta, then H_{k+1} applied over
    // (mode, beta) gives the next Y (or the term's part of the last core)
    const int64_t n1 = sp.dims[k + 1];
    const bool last = k + 2 == d;
    std::vector<char> has_op(items.size(), 0);
    for (size_t i = 0; i < items.size(); ++i) {
      const int64_t r1 = orr[i][k + 1];
      for (size_t tt = 0; tt < items[i].terms.size(); ++tt) {
        const RandTerm& t = items[i].terms[tt];
        if (!t.H) continue;
        has_op[i] = 1;
        const int64_t ry1 = t.x->rks[k + 1], ry2 = t.x->rks[k + 2], Hl = t.H->rks[k + 1];
        const double* Pb = Pat(i, tt, k);
        for (int64_t b = 0; b < Hl; ++b)
          gb.add(int(r1), int(n1 * ry2), int(ry1), 1.0, Pb + r1 * ry1 * b, 1, r1, t.x->cores[k + 1], 1, ry1, 0.0,
                 t.U + r1 * n1 * ry2 * b, 1, r1);
        OpApply a = op_job(t.H, k + 1, false);
        a.in = t.U;
        a.out = last ? outs[i]->cores[k + 1] : Yat(i, tt, k + 1, yb ^ 1);
        a.X = int(r1);
        a.Y = int(ry2);
        a.is_m = r1; a.is_b = r1 * n1 * ry2; a.is_x = 1; a.is_y = r1 * n1;
        a.os_m = r1; a.os_b = r1 * n1 * ry2; a.os_x = 1; a.os_y = r1 * n1;
        oj.push_back(a);
      }
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "rr_U");
    op_apply_batched(ctx, oj, Prec::F64, Prec::F64);
    oj.clear();
    for (size_t i = 0; i < items.size(); ++i) {
      const int64_t r1 = orr[i][k + 1];
      if (!last) {
        for (size_t tt = 0; tt < items[i].terms.size(); ++tt) {
          const RandTerm& t = items[i].terms[tt];
          if (t.H) continue;
          const int64_t R1 = t.x->rks[k + 1], R2 = t.x->rks[k + 2];
          gb.add(int(r1), int(n1 * R2), int(R1), 1.0, Pat(i, tt, k), 1, r1, t.x->cores[k + 1], 1, R1, 0.0,
                 Yat(i, tt, k + 1, yb ^ 1), 1, r1);
        }
      } else {
        bool any = false;
        for (size_t tt = 0; tt < items[i].terms.size(); ++tt) {
          const RandTerm& t = items[i].terms[tt];
          if (t.H) continue;
          if (!any) gb.prob(outs[i]->cores[k + 1], 1, r1, int(r1), int(n1), 1.0, has_op[i] ? 1.0 : 0.0);
          any = true;
          const int64_t R1 = t.x->rks[k + 1];
          gb.seg(Pat(i, tt, k), 1, r1, t.x->cores[k + 1], 1, R1, int(R1), 1.0);
        }
      }
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "rr_Y");
    yb ^= 1;
  }
  return outs;
}

std::vector<RandItem> parse_items(ttr_handle_s* h, int64_t n, const ttr_rand_item* items) {
  TTR_REQUIRE(n >= 1 && items != nullptr, ST_INVALID_ARG, "need at least one item");
  std::vector<RandItem> out(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    const ttr_rand_item& it = items[i];
    TTR_REQUIRE(it.nterms >= 1 && it.term_ids && it.alpha && it.rks, ST_INVALID_ARG, "bad randomized item");
    for (int64_t t = 0; t < it.nterms; ++t) {
      RandTerm rt{};
      rt.x = h->get(it.term_ids[t]);
      rt.alpha = it.alpha[t];
      if (t == 0 && it.op_id >= 0) rt.H = h->get_op(it.op_id);
      out[i].terms.push_back(rt);
    }
    const int d = out[i].terms[0].x->d;
    out[i].rks.assign(it.rks, it.rks + d + 1);
  }
  return out;
}

// Key of a ttr_round_rand call: what fixes its launches and its workspace layout
// (shapes, ranks, options, seed).  Coefficients, TT contents and TT addresses are
// not part of it: they reach the kernels only through uploaded descriptors.
std::string rand_call_key(const std::vector<RandItem>& items, const ttr_rand_opts* o) {
  std::string k;
  auto put = [&k](int64_t v) { k.append(reinterpret_cast<const char*>(&v), sizeof(v)); };
  put(int64_t(items.size()));
  for (const auto& it : items) {
    put(int64_t(it.terms.size()));
    for (const auto& t : it.terms) {
      put(t.x->d);
      for (int64_t v : t.x->dims) put(v);
      for (int64_t v : t.x->rks) put(v);
      put(t.H ? 1 : 0);
      if (t.H)
        for (int64_t v : t.H->rks) put(v);
    }
    for (int64_t v : it.rks) put(v);
  }
  int64_t ov;
  std::memcpy(&ov, &o->oversampling, sizeof(ov));
  put(o->orthogonal);
  put(o->sketch_prec);
  put(o->block_rks);
  put(int64_t(o->seed));
  put(ov);
  put(o->forward);
  put(o->sketch_accum);
  put(o->blocks ? 1 : 0);
  put(o->p ? 1 : 0);
  if (o->p)
    for (int q = 0; q <= items[0].terms[0].x->d; ++q) put(o->p[q]);
  return k;
}

// ttr_round_rand through HIP graphs (handle option, off by default).  Per call
// key: the first call runs eagerly and records how much workspace and pinned
// staging it used; the second is captured (stream capture, relaxed mode) into a
// graph whose workspace and staging are fixed regions owned by the cache entry,
// then launched; later calls rerun the host side only, which writes the call's
// descriptors (TT core addresses, coefficients, GEMM problems) into the same
// pinned staging that the graph's copy nodes read, and launch the graph.  Such a
// replay must record the same uploads (sizes and addresses) and the same
// launches (kernel, grid, block and argument bytes) as the capture; if it does
// not, the call runs eagerly and the key is not graphed again.  The split-K
// partial sums take one region of the fixed workspace sized for the largest
// request of the first call, so every GEMM splits K as in the eager call and the
// results are the eager ones bit for bit.
std::vector<std::unique_ptr<DevTT>> round_rand_graphed(ttr_handle_s* h, std::vector<RandItem>& items,
                                                       const ttr_rand_opts* o) {
  Ctx& ctx = h->ctx;
  if (!h->graphs || ctx.prof.on || !o) return round_rand_batch(h, items, o);
  const std::string key = rand_call_key(items, o);
  auto found = h->gcache.find(key);
  if (found == h->gcache.end()) {
    constexpr size_t kMaxGraphs = 8;
    while (h->gcache.size() >= kMaxGraphs) {  // evict the least recently used shape
      auto lru = h->gcache.begin();
      for (auto it = h->gcache.begin(); it != h->gcache.end(); ++it)
        if (it->second.last_use < lru->second.last_use) lru = it;
      TTR_HIP(hipStreamSynchronize(ctx.stream));
      ttr_handle_s::graph_free(lru->second);
      h->gcache.erase(lru);
    }
    found = h->gcache.emplace(key, ttr_handle_s::GraphEntry{}).first;
  }
  ttr_handle_s::GraphEntry& e = found->second;
  e.last_use = ++h->gclock;
  if (e.state < 0) return round_rand_batch(h, items, o);
  if (e.state == 0 && e.ws_need == 0) {
    ctx.splitk_peak = 0;
    // uploads as in the capture, so the measured workspace fits it
    struct UploadAll {
      Ctx& c;
      explicit UploadAll(Ctx& cx) : c(cx) { c.upload_all = true; }
      ~UploadAll() { c.upload_all = false; }
    } upload_all(ctx);
    auto outs = round_rand_batch(h, items, o);
    e.ws_need = ctx.ws.used() + ((ctx.splitk_peak * sizeof(double) + 255) & ~size_t(255)) + 256;
    e.pin_need = ctx.pin.used() + 256;
    e.split_elems = ctx.splitk_peak;
    return outs;
  }
  // fixed regions, graph mode and capture are undone whatever happens
  struct Scope {
    Ctx& c;
    bool capturing = false;
    explicit Scope(Ctx& cx) : c(cx) {}
    void end() {
      if (capturing) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        capturing = false;
      }
      c.ws.clear_fixed();
      c.pin.clear_fixed();
      c.fsplit = nullptr;
      c.fsplit_elems = 0;
      c.gmode = GraphMode::Off;
    }
    ~Scope() { end(); }
  } scope(ctx);
  auto enter = [&](GraphMode m) {
    ctx.ws.set_fixed(e.dev, e.dev_cap);
    ctx.pin.set_fixed(e.host, e.host_cap);
    ctx.fsplit = e.split_elems ? ctx.ws.alloc_n<double>(e.split_elems) : nullptr;
    ctx.fsplit_elems = e.split_elems;
    ctx.gsig.clear();
    ctx.gmode = m;
  };
  auto eager_again = [&](std::vector<std::unique_ptr<DevTT>>& outs) {
    scope.end();
    for (auto& t : outs)
      if (t) h->free_tt(t.get());
    e.state = -1;
    ttr_handle_s::graph_free(e);
    ++h->g_fallbacks;
    return round_rand_batch(h, items, o);
  };
  if (e.state == 0) {
    e.dev_cap = e.ws_need;
    e.host_cap = e.pin_need;
    TTR_HIP(hipMalloc(&e.dev, e.dev_cap));
    TTR_HIP(hipHostMalloc(&e.host, e.host_cap, hipHostMallocDefault));
    enter(GraphMode::Capture);
    TTR_HIP(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeRelaxed));
    scope.capturing = true;
    std::vector<std::unique_ptr<DevTT>> outs;
    hipGraph_t g = nullptr;
    hipError_t st = hipSuccess;
    try {
      outs = round_rand_batch(h, items, o);
    } catch (const Error&) {
      st = hipErrorUnknown;
    }
    scope.capturing = false;
    const hipError_t ec = hipStreamEndCapture(ctx.stream, &g);
    if (st == hipSuccess) st = ec;
    hipGraphExec_t ex = nullptr;
    if (st == hipSuccess) st = hipGraphInstantiateWithFlags(&ex, g, 0);
    if (g) (void)hipGraphDestroy(g);
    if (st != hipSuccess) {
      (void)hipGetLastError();
      return eager_again(outs);
    }
    e.exec = ex;
    e.sig = ctx.gsig;
    e.state = 1;
    TTR_HIP(hipGraphLaunch(e.exec, ctx.stream));
    ++h->g_captures;
    return outs;
  }
  enter(GraphMode::Replay);
  auto outs = round_rand_batch(h, items, o);
  if (ctx.gsig != e.sig) return eager_again(outs);
  TTR_HIP(hipGraphLaunch(e.exec, ctx.stream));
  ++h->g_replays;
  return outs;
}

// ------------------------------------------------------------ forward sketch
// Block ranks of the forward (left-to-right) sketch (src/sketches.jl:244-248):
// bv_0 = 1, bv_{k+1} = min(block_rks, n_k bv_k)
std::vector<int64_t> block_ranks_fwd(const std::vector<int64_t>& dims, int64_t block_rks) {
  const int d = int(dims.size());
  std::vector<int64_t> bv(d + 1, 1);
  for (int k = 1; k <= d; ++k) bv[k] = block_rks;
  for (int k = 0; k < d; ++k) bv[k + 1] = std::min(bv[k + 1], dims[k] * bv[k]);
  return bv;
}

// compute_sketch_blocks_heuristic(...; reverse=false) (src/sketches.jl:113-121) and the
// oversampling of p_2..p_{N+1} (:253-255)
std::vector<int64_t> block_counts_fwd(const std::vector<int64_t>& rks, const std::vector<int64_t>& bv, int d,
                                      const int64_t* p_in, double oversampling) {
  std::vector<int64_t> p(d + 1, 1);
  if (p_in) {
    for (int k = 0; k <= d; ++k) p[k] = p_in[k];
  } else {
    for (int k = 0; k <= d; ++k) p[k] = int64_t(std::ceil(double(rks[k]) / double(bv[k])));
    for (int k = d - 1; k >= 1; --k) p[k] = std::max(p[k], p[k + 1]);
    p[0] = 1;
  }
  if (oversampling != 1.0 && oversampling > 0.0)
    for (int k = 1; k <= d; ++k) p[k] = int64_t(std::ceil(oversampling * double(p[k])));
  for (int k = 0; k <= d; ++k) TTR_REQUIRE(p[k] >= 1, ST_INVALID_ARG, "block counts must be >= 1");
  TTR_REQUIRE(p[0] == 1, ST_INVALID_ARG, "the first block count of a forward sketch must be 1");
  for (int k = 1; k < d; ++k)
    TTR_REQUIRE(p[k] >= p[k + 1], ST_INVALID_ARG, "forward block counts must be non-increasing towards the right");
  return p;
}

struct FwdPlan {
  int d = 0;
  std::vector<int64_t> dims, bv, pmax;
  bool orthogonal = true;
  uint64_t seed = 1234;
  const double* const* blocks = nullptr;  // caller-supplied blocks (reference layout), or null
  std::vector<double*> S;  // core k: (bv_k n_k) x bv_{k+1} x pmax_{k+1}, reference order, FP64
};

// One term of a forward sketch: W_k (r_k x bv_k p_k), unscaled (the 1/sqrt(p_k) is
// applied where W_k is used), k = 1..kend; W_0 = 1
struct FwdTerm {
  DevTT* x = nullptr;
  DevTTO* H = nullptr;     // operator term: the sketch of H x (never formed)
  double alpha = 1.0;
  std::vector<int64_t> p;  // block counts of its item
  std::vector<double*> W;  // W_k: (rc(k) x bv_k p_k), rows (x rank, H rank) with x rank fastest
  int64_t rc(int k) const { return x->rks[k] * (H ? H->rks[k] : 1); }
};

// generate_sketch_blocks(rng, T, bv_k, n_k, bv_{k+1}, p_{k+1}, orthogonal) for cores
// 0..kend-1 (src/sketches.jl:44-84,279-281): Gaussian blocks scaled by 1/sqrt(bv_{k+1});
// orthogonal: thin Householder Q of each block scaled by sqrt(bv_k n_k / bv_{k+1}), or
// identity blocks when bv_{k+1} = n_k bv_k
void make_sketch_fwd(ttr_handle_s* h, FwdPlan& fp, int kend) {
  Ctx& ctx = h->ctx;
  fp.S.assign(fp.d, nullptr);
  if (fp.blocks) {  // the caller's blocks are in this layout already
    for (int k = 0; k < kend; ++k) {
      TTR_REQUIRE(fp.blocks[k] != nullptr, ST_INVALID_ARG, "blocks[" + std::to_string(k) + "] is NULL");
      const int64_t cnt = fp.bv[k] * fp.dims[k] * fp.bv[k + 1] * fp.pmax[k + 1];
      fp.S[k] = const_cast<double*>(ctx.upload(fp.blocks[k], size_t(cnt)));
    }
    return;
  }
  std::vector<SketchGen> gen;
  std::vector<Fill2D> fills;
  std::vector<QRJob> qj;
  std::vector<Copy2D> cp;
  for (int k = 0; k < kend; ++k) {
    const int64_t n = fp.dims[k], left = fp.bv[k], right = fp.bv[k + 1], p = fp.pmax[k + 1], m = left * n;
    fp.S[k] = ctx.ws.alloc_n<double>(size_t(m * right * p));
    if (!fp.orthogonal) {
      SketchGen g{nullptr, fp.S[k], int(n), int(left), int(right), int(p), k, 1.0};
      g.ref_scale = 1.0 / std::sqrt(double(right));
      gen.push_back(g);
    } else if (right == m) {
      for (int64_t j = 0; j < p; ++j) fills.push_back(Fill2D{fp.S[k] + j * m * right, m, int(m), int(right), 1.0});
    } else {
      double* raw = ctx.ws.alloc_n<double>(size_t(m * right * p));
      double* q = ctx.ws.alloc_n<double>(size_t(m * right * p));
      gen.push_back(SketchGen{nullptr, raw, int(n), int(left), int(right), int(p), k, 1.0});
      for (int64_t j = 0; j < p; ++j) {
        qj.push_back(QRJob{raw + j * m * right, m, int(m), int(right), q + j * m * right, m});
        cp.push_back(copy_job(q + j * m * right, 1, m, fp.S[k] + j * m * right, 1, m, int(m), int(right),
                              std::sqrt(double(m) / double(right))));
      }
    }
  }
  fill2d_batched(ctx, fills);
  sketch_generate(ctx, gen, fp.seed, Prec::F64);
  qr_batched(ctx, qj);
  copy2d_batched(ctx, cp);
}

// forward recursive sketch of every term (src/sketches.jl:276-298), all terms in the same
// launches: for core k and block j, T1 = W_k[:, block j]^T X_k (bv_k x n r_{k+1}), then
// W_{k+1}[:, block j] = T1^T S_j with T1 read as (bv_k n) x r_{k+1}; W_0 = 1 serves every
// block of core 0
void sketch_phase_fwd(ttr_handle_s* h, const FwdPlan& fp, std::vector<FwdTerm>& terms, int kend) {
  Ctx& ctx = h->ctx;
  GemmBatch gb;
  for (auto& t : terms) {
    t.W.assign(size_t(fp.d) + 1, nullptr);
    for (int k = 1; k <= kend; ++k) t.W[k] = ctx.ws.alloc_n<double>(size_t(t.rc(k) * fp.bv[k] * t.p[k]));
  }
  for (int k = 0; k < kend; ++k) {
    const int64_t n = fp.dims[k], bl = fp.bv[k], br = fp.bv[k + 1];
    std::vector<double*> T1(terms.size(), nullptr), T2(terms.size(), nullptr);
    if (k > 0) {
      // T1 = W_k[:, block j]^T X_k summed over the x rank: rows (H rank, bl), columns (n, r1)
      for (size_t i = 0; i < terms.size(); ++i) {
        const DevTT* x = terms[i].x;
        const int64_t r = x->rks[k], r1 = x->rks[k + 1], pk1 = terms[i].p[k + 1];
        const int64_t R = terms[i].H ? terms[i].H->rks[k] : 1;
        T1[i] = ctx.ws.alloc_n<double>(size_t(R * bl * n * r1 * pk1));
        for (int64_t j = 0; j < pk1; ++j)
          gb.add(int(R * bl), int(n * r1), int(r), 1.0, terms[i].W[k] + r * R * bl * j, r, 1, x->cores[k], 1, r, 0.0,
                 T1[i] + R * bl * n * r1 * j, 1, R * bl);
      }
      gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "fwd_t1");
    }
    // operator terms (contract_sketch_core_forwards!, src/sketches.jl:1240-1309): the
    // operator core acts on the (mode, left H bond) pairs of T1,
    // T2[(g, z), (a, b)] = sum_{zeta, beta} H[z, zeta, beta, b] T1[(beta, g), (zeta, a)]
    std::vector<OpApply> oj;
    for (size_t i = 0; i < terms.size(); ++i) {
      const DevTTO* H = terms[i].H;
      if (!H) continue;
      const DevTT* x = terms[i].x;
      TTR_REQUIRE(!H->wide(k), ST_INVALID_ARG, "the forward operator sketch needs n_k * R_k <= 64");
      const int64_t R = H->rks[k], R1 = H->rks[k + 1], r1 = x->rks[k + 1];
      const int64_t nb = k == 0 ? 1 : terms[i].p[k + 1];  // W_0 = 1 serves every block of core 0
      T2[i] = ctx.ws.alloc_n<double>(size_t(bl * n * r1 * R1 * nb));
      for (int64_t j = 0; j < nb; ++j) {
        OpApply o = op_job(H, k, false);
        o.in = k == 0 ? static_cast<const double*>(x->cores[0]) : T1[i] + R * bl * n * r1 * j;
        o.out = T2[i] + bl * n * r1 * R1 * j;
        o.X = int(bl);
        o.Y = int(r1);
        o.is_b = 1; o.is_x = R; o.is_m = R * bl; o.is_y = R * bl * n;
        o.os_x = 1; o.os_m = bl; o.os_y = bl * n; o.os_b = bl * n * r1;
        oj.push_back(o);
      }
    }
    op_apply_batched(ctx, oj, Prec::F64, Prec::F64);
    for (size_t i = 0; i < terms.size(); ++i) {
      const DevTT* x = terms[i].x;
      const int64_t r1 = x->rks[k + 1], pk1 = terms[i].p[k + 1];
      if (terms[i].H) {  // W_{k+1}[(a, b), c] = T2^T S_j over (g, z)
        const int64_t R1 = terms[i].H->rks[k + 1];
        for (int64_t j = 0; j < pk1; ++j)
          gb.add(int(r1 * R1), int(br), int(bl * n), 1.0, T2[i] + (k == 0 ? 0 : bl * n * r1 * R1 * j), bl * n, 1,
                 fp.S[k] + bl * n * br * j, 1, bl * n, 0.0, terms[i].W[k + 1] + r1 * R1 * br * j, 1, r1 * R1);
        continue;
      }
      for (int64_t j = 0; j < pk1; ++j) {
        const double* A = k == 0 ? x->cores[0] : T1[i] + bl * n * r1 * j;  // (bl n) x r1
        gb.add(int(r1), int(br), int(bl * n), 1.0, A, bl * n, 1, fp.S[k] + bl * n * br * j, 1, bl * n, 0.0,
               terms[i].W[k + 1] + r1 * br * j, 1, r1);
      }
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "fwd_w");
  }
}

void prepare_fwd(const ttr_rand_opts* o, int d, const std::vector<int64_t>& dims, FwdPlan& fp) {
  fp.d = d;
  fp.dims = dims;
  fp.orthogonal = o->orthogonal != 0;
  fp.seed = o->seed;
  fp.blocks = o->blocks;
  fp.bv = block_ranks_fwd(dims, o->block_rks > 0 ? o->block_rks : d);
  fp.pmax.assign(size_t(d) + 1, 1);
}

// ------------------------------------------------------------ STTA
// One item of stta(alpha, y, rks) (src/tt_randtools.jl:552-638): the sketches of every
// term, then Omega_k = sum_t alpha_t W_L,t[k+1]^T W_R,t[k+1] and the sketched cores
// Psi_k = sum_t alpha_t X_t,k x_1 W_L,t[k] x_3 W_R,t[k+1] (device layout (sl_k, n_k, sr_{k+1}))
struct SttaItem {
  std::vector<FwdTerm> left;  // per term
  RandItem right;             // reverse sketch of the same terms
  std::vector<int64_t> pl, sl, sr;
  std::vector<double*> Om;    // d-1: sl_{k+1} x sr_{k+1}
  std::vector<double*> Psi;   // d: (sl_k, n_k, sr_{k+1})
};

void stta_sketch_batch(ttr_handle_s* h, std::vector<SttaItem>& its, const ttr_rand_opts* o, uint64_t seed_right) {
  Ctx& ctx = h->ctx;
  TTR_REQUIRE(o != nullptr, ST_INVALID_ARG, "options must not be NULL");
  TTR_REQUIRE(to_prec(o->sketch_prec) == Prec::F64, ST_INVALID_ARG, "STTA sketches are FP64");
  TTR_REQUIRE(o->p == nullptr, ST_INVALID_ARG, "STTA takes its block counts from the ranks");
  const DevTT* x0 = its[0].right.terms[0].x;
  const int d = x0->d;
  for (auto& it : its)
    for (auto& t : it.right.terms) {
      TTR_REQUIRE(t.x->dims == x0->dims, ST_DIM_MISMATCH, "all terms must share dims");
      TTR_REQUIRE(t.x->rks[0] == 1 && t.x->rks[d] == 1, ST_INVALID_ARG, "boundary ranks must be 1");
      TTR_REQUIRE(t.H == nullptr, ST_INVALID_ARG, "STTA takes no operator");
    }
  // right: the reverse sketch (no oversampling), seed_right
  ttr_rand_opts orr = *o;
  orr.seed = seed_right;
  orr.oversampling = 1.0;
  orr.forward = 0;
  orr.blocks = o->blocks_right;
  TTR_REQUIRE((o->blocks == nullptr) == (o->blocks_right == nullptr), ST_INVALID_ARG,
              "STTA takes caller-supplied blocks for both sketches or for neither");
  std::vector<RandItem> ritems;
  for (auto& it : its) ritems.push_back(it.right);
  SketchPlan sp;
  prepare_plan(h, &orr, d, x0->dims, ritems, sp);
  // left: the forward sketch with oversampling 1.5 (src/sketches.jl:1131-1133), seed_left
  FwdPlan fp;
  prepare_fwd(o, d, x0->dims, fp);
  std::vector<FwdTerm> all;
  for (auto& it : its) {
    it.pl = block_counts_fwd(it.right.rks, fp.bv, d, nullptr, 1.5);
    it.sl.resize(size_t(d) + 1);
    for (int k = 0; k <= d; ++k) {
      it.sl[k] = fp.bv[k] * it.pl[k];
      fp.pmax[k] = std::max(fp.pmax[k], it.pl[k]);
    }
    TTR_REQUIRE(!fp.blocks || it.pl == its[0].pl, ST_INVALID_ARG,
                "with caller-supplied blocks every item must have the same block counts");
  }
  {
    ProfScope ps(ctx, "phase:stta_sketch");
    make_sketch_fwd(h, fp, d - 1);
    for (auto& it : its)
      for (auto& t : it.right.terms) {
        FwdTerm ft;
        ft.x = t.x;
        ft.alpha = t.alpha;
        ft.p = it.pl;
        all.push_back(ft);
      }
    sketch_phase_fwd(h, fp, all, d - 1);
    make_sketch(h, sp, 1);
    sketch_phase(h, sp, ritems, 1);
  }
  size_t ti = 0;
  for (size_t i = 0; i < its.size(); ++i) {
    its[i].right = ritems[i];
    its[i].sr = ritems[i].s;
    its[i].left.assign(all.begin() + ti, all.begin() + ti + ritems[i].terms.size());
    ti += ritems[i].terms.size();
  }
  // the power-of-two exponents of the right sketches (the GEMM's per-segment exponents are
  // relative to the problem's largest, so the absolute factors go into the host scales)
  std::vector<int32_t*> eh;
  for (auto& it : its)
    for (auto& t : it.right.terms) {
      int32_t* e = static_cast<int32_t*>(ctx.pin.alloc(sizeof(int32_t) * (size_t(d) + 1)));
      TTR_HIP(hipMemcpyAsync(e, t.E, sizeof(int32_t) * (size_t(d) + 1), hipMemcpyDeviceToHost, ctx.stream));
      eh.push_back(e);
    }
  ctx.sync();
  std::vector<std::vector<const int32_t*>> E(its.size());
  ti = 0;
  for (size_t i = 0; i < its.size(); ++i)
    for (size_t t = 0; t < its[i].right.terms.size(); ++t) E[i].push_back(eh[ti++]);
  GemmBatch gb;
  // Omega_k, k = 0..d-2
  for (size_t i = 0; i < its.size(); ++i) {
    auto& it = its[i];
    it.Om.assign(size_t(std::max(d - 1, 0)), nullptr);
    for (int k = 0; k + 1 < d; ++k) {
      const int64_t sl = it.sl[k + 1], sr = it.sr[k + 1];
      it.Om[k] = ctx.ws.alloc_n<double>(size_t(sl * sr));
      gb.prob(it.Om[k], 1, sl, int(sl), int(sr), 1.0, 0.0);
      for (size_t t = 0; t < it.left.size(); ++t) {
        const RandTerm& rt = it.right.terms[t];
        const int64_t r = rt.x->rks[k + 1];
        const double sc = std::ldexp(it.left[t].alpha, E[i][t][k + 1]) /
                          std::sqrt(double(it.pl[k + 1]) * double(it.right.p[k + 1]));
        gb.seg(it.left[t].W[k + 1], r, 1, rt.W[k + 1], 1, r, int(r), sc);
      }
    }
  }
  gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "stta_omega");
  // T_k,t = W_L,t[k]^T X_t,k (sl_k x n r_{k+1}) for k >= 1
  std::vector<std::vector<std::vector<double*>>> T(its.size());
  for (size_t i = 0; i < its.size(); ++i) {
    auto& it = its[i];
    T[i].assign(it.left.size(), std::vector<double*>(size_t(d), nullptr));
    for (size_t t = 0; t < it.left.size(); ++t) {
      const DevTT* x = it.left[t].x;
      for (int k = 1; k < d; ++k) {
        const int64_t sl = it.sl[k], r = x->rks[k], n = x->dims[k], r1 = x->rks[k + 1];
        T[i][t][k] = ctx.ws.alloc_n<double>(size_t(sl * n * r1));
        gb.add(int(sl), int(n * r1), int(r), 1.0, it.left[t].W[k], r, 1, x->cores[k], 1, r, 0.0, T[i][t][k], 1, sl);
      }
    }
  }
  gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "stta_t");
  // Psi_k = sum_t alpha_t T_k,t W_R,t[k+1] / sqrt(pl_k pr_{k+1}); W_L[0] = W_R[d] = 1
  for (size_t i = 0; i < its.size(); ++i) {
    auto& it = its[i];
    it.Psi.assign(size_t(d), nullptr);
    for (int k = 0; k < d; ++k) {
      const int64_t sl = it.sl[k], n = x0->dims[k], sr = it.sr[k + 1];
      it.Psi[k] = ctx.ws.alloc_n<double>(size_t(sl * n * sr));
      gb.prob(it.Psi[k], 1, sl * n, int(sl * n), int(sr), 1.0, 0.0);
      for (size_t t = 0; t < it.left.size(); ++t) {
        const RandTerm& rt = it.right.terms[t];
        const int64_t r1 = rt.x->rks[k + 1];
        const double* A = k == 0 ? rt.x->cores[0] : T[i][t][k];
        const double sc = it.left[t].alpha / std::sqrt(double(it.pl[k]) * double(it.right.p[k + 1]));
        if (k + 1 < d) gb.seg(A, 1, sl * n, rt.W[k + 1], 1, r1, int(r1), std::ldexp(sc, E[i][t][k + 1]));
        else gb.seg(A, 1, sl * n, h->one, 1, 1, 1, sc);
      }
    }
  }
  gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "stta_psi");
}

// stable_inverse(A; eps = 1e-12) (src/tt_randtools.jl:485-488) of every Omega_k and the
// local solves of stta (:562-575): per bond, Psi_k <- Psi_k pinv(Omega_k) when Omega_k has
// fewer rows than columns, else Psi_{k+1} <- pinv(Omega_k) Psi_{k+1}.  pinv of an m x nn
// matrix M (m >= nn; Omega_k or its transpose): M = Q R, R J = U S (Jacobi), pinv(M) =
// J diag(1/s^2) (Q R J)^T over s > eps max(s).  The left products act on the left index
// and the right products on the right index of a core, so all of them are batched.
std::vector<std::unique_ptr<DevTT>> stta_solve_batch(ttr_handle_s* h, std::vector<SttaItem>& its) {
  Ctx& ctx = h->ctx;
  const int d = its[0].right.terms[0].x->d;
  const std::vector<int64_t>& dims = its[0].right.terms[0].x->dims;
  struct Bond {
    bool wide;      // sl < sr: Psi_k <- Psi_k pinv
    int64_t m, nn;  // M = m x nn
    double *M, *Q, *B, *J, *sv, *QB, *P;  // P = pinv(M), nn x m
  };
  std::vector<std::vector<Bond>> bonds(its.size());
  std::vector<Copy2D> cp;
  std::vector<QRJob> qj;
  for (size_t i = 0; i < its.size(); ++i)
    for (int k = 0; k + 1 < d; ++k) {
      const int64_t sl = its[i].sl[k + 1], sr = its[i].sr[k + 1];
      Bond b{};
      b.wide = sl < sr;
      b.m = b.wide ? sr : sl;
      b.nn = b.wide ? sl : sr;
      b.M = ctx.ws.alloc_n<double>(size_t(b.m * b.nn));
      b.Q = ctx.ws.alloc_n<double>(size_t(b.m * b.nn));
      b.B = ctx.ws.alloc_n<double>(size_t(b.nn * b.nn));
      b.J = ctx.ws.alloc_n<double>(size_t(b.nn * b.nn));
      b.sv = ctx.ws.alloc_n<double>(size_t(b.nn));
      b.QB = ctx.ws.alloc_n<double>(size_t(b.m * b.nn));
      b.P = ctx.ws.alloc_n<double>(size_t(b.nn * b.m));
      if (b.wide) cp.push_back(copy_job(its[i].Om[k], sl, 1, b.M, 1, b.m, int(b.m), int(b.nn)));
      else cp.push_back(copy_job(its[i].Om[k], 1, sl, b.M, 1, b.m, int(b.m), int(b.nn)));
      qj.push_back(QRJob{b.M, b.m, int(b.m), int(b.nn), b.Q, b.m});
      bonds[i].push_back(b);
    }
  copy2d_batched(ctx, cp);
  qr_batched(ctx, qj);
  cp.clear();
  std::vector<SVDJob> sj;
  for (auto& bb : bonds)
    for (auto& b : bb) {
      cp.push_back(copy_job(b.M, 1, b.m, b.B, 1, b.nn, int(b.nn), int(b.nn), 1.0, 1));
      sj.push_back(SVDJob{b.B, b.nn, b.J, b.nn, b.sv, int(b.nn), nullptr});
    }
  copy2d_batched(ctx, cp);
  if (!sj.empty()) jacobi_svd_batched(ctx, sj);
  std::vector<PinvWeights> pw;
  for (auto& bb : bonds)
    for (auto& b : bb) pw.push_back(PinvWeights{b.B, b.nn, b.sv, int(b.nn), 1e-12});
  pinv_weights_batched(ctx, pw);
  GemmBatch gb;
  for (auto& bb : bonds)
    for (auto& b : bb) gb.add(int(b.m), int(b.nn), int(b.nn), 1.0, b.Q, 1, b.m, b.B, 1, b.nn, 0.0, b.QB, 1, b.m);
  gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "stta_pinv_a");
  for (auto& bb : bonds)
    for (auto& b : bb) gb.add(int(b.nn), int(b.m), int(b.nn), 1.0, b.J, 1, b.nn, b.QB, b.m, 1, 0.0, b.P, 1, b.nn);
  gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "stta_pinv_b");
  // output ranks: bond k keeps sl_{k+1} (wide) or sr_{k+1}
  std::vector<std::unique_ptr<DevTT>> outs(its.size());
  std::vector<std::vector<double*>> mid(its.size());
  for (size_t i = 0; i < its.size(); ++i) {
    std::vector<int64_t> r(size_t(d) + 1, 1);
    for (int k = 0; k + 1 < d; ++k) r[k + 1] = bonds[i][k].wide ? its[i].sl[k + 1] : its[i].sr[k + 1];
    outs[i] = h->make_tt(dims, r, std::vector<int64_t>(size_t(d), 0));
    mid[i].assign(size_t(d), nullptr);
  }
  // left products: core k+1 <- pinv_k Psi_{k+1} for the tall bonds (into the output core
  // when the core takes no right product, else into a scratch core)
  for (size_t i = 0; i < its.size(); ++i)
    for (int k = 0; k + 1 < d; ++k) {
      const Bond& b = bonds[i][k];
      if (b.wide) continue;
      const int c = k + 1;
      const int64_t sl = its[i].sl[c], n = dims[c], sr = its[i].sr[c + 1];
      const bool right_next = c + 1 < d && bonds[i][c].wide;
      double* dst = right_next ? ctx.ws.alloc_n<double>(size_t(b.nn * n * sr)) : outs[i]->cores[c];
      mid[i][c] = dst;
      gb.add(int(b.nn), int(n * sr), int(sl), 1.0, b.P, 1, b.nn, its[i].Psi[c], 1, sl, 0.0, dst, 1, b.nn);
    }
  gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "stta_left");
  // right products: core k <- core k pinv_k = core k P^T for the wide bonds; cores with
  // neither product are copied
  cp.clear();
  for (size_t i = 0; i < its.size(); ++i)
    for (int c = 0; c < d; ++c) {
      const bool left = c > 0 && !bonds[i][c - 1].wide;
      const bool right = c + 1 < d && bonds[i][c].wide;
      const int64_t lr = outs[i]->rks[c], n = dims[c];
      const double* src = left ? mid[i][c] : its[i].Psi[c];
      if (right) {
        const Bond& b = bonds[i][c];
        gb.add(int(lr * n), int(b.nn), int(b.m), 1.0, src, 1, lr * n, b.P, b.nn, 1, 0.0, outs[i]->cores[c], 1, lr * n);
      } else if (!left) {
        const int64_t rr = outs[i]->rks[c + 1];
        cp.push_back(copy_job(src, 1, lr * n, outs[i]->cores[c], 1, lr * n, int(lr * n), int(rr)));
      }
    }
  gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "stta_right");
  copy2d_batched(ctx, cp);
  return outs;
}

std::vector<SttaItem> parse_stta_items(ttr_handle_s* h, int64_t n, const ttr_rand_item* items) {
  std::vector<RandItem> ri = parse_items(h, n, items);
  std::vector<SttaItem> out(ri.size());
  for (size_t i = 0; i < ri.size(); ++i) {
    TTR_REQUIRE(items[i].op_id < 0, ST_INVALID_ARG, "STTA takes no operator");
    out[i].right = ri[i];
    TTR_REQUIRE(out[i].right.rks.size() == out[i].right.terms[0].x->dims.size() + 1, ST_INVALID_ARG, "bad rks");
    TTR_REQUIRE(out[i].right.rks.front() == 1 && out[i].right.rks.back() == 1, ST_INVALID_ARG,
                "rks must have rks[0] == rks[d] == 1");
  }
  return out;
}

// ------------------------------------------------------------------ C-ABI glue
template <class F>
int guarded(ttr_handle_t h, F&& f) {
  if (!h) return TTR_INVALID_ARG;
  try {
    h->err.clear();
    f();
    h->ctx.end_call();
    return TTR_OK;
  } catch (const Error& e) {
    h->err = e.what();
    (void)hipGetLastError();
    (void)hipStreamSynchronize(h->ctx.stream);
    h->ctx.ws.reset();
    h->ctx.pin.reset();
    return e.status;
  } catch (const std::exception& e) {
    h->err = e.what();
    (void)hipStreamSynchronize(h->ctx.stream);
    h->ctx.ws.reset();
    h->ctx.pin.reset();
    return TTR_DEVICE_ERROR;
  }
}

}  // namespace

extern "C" {

const char* ttr_version(void) { return "libttround 0.1.0 (gfx950)"; }

int ttr_create(ttr_handle_t* out, int device) {
  if (!out) return TTR_INVALID_ARG;
  *out = nullptr;
  ttr_handle_s* h = new (std::nothrow) ttr_handle_s();
  if (!h) return TTR_OOM;
  try {
    int count = 0;
    TTR_HIP(hipGetDeviceCount(&count));
    TTR_REQUIRE(device >= 0 && device < count, ST_INVALID_ARG, "no such HIP device");
    TTR_HIP(hipSetDevice(device));
    h->ctx.device = device;
    TTR_HIP(hipStreamCreateWithFlags(&h->ctx.stream, hipStreamNonBlocking));
    TTR_HIP(hipMalloc(&h->one, 64));
    h->scratch = scratch_for_device(device);
    TTR_HIP(hipHostMalloc(&h->stage, TTR_STAGE_BYTES, hipHostMallocDefault));
    const double one = 1.0;
    TTR_HIP(hipMemcpy(h->one, &one, 8, hipMemcpyHostToDevice));
    // TTR_GEMM_TUNE=0: GEMM tiles by size only (no timing of the candidates on first use)
    const char* tune = std::getenv("TTR_GEMM_TUNE");
    h->ctx.gemm_autotune = !(tune && tune[0] == '0');
    // TTR_GEMM_CFG=c (0..7): every GEMM uses tile configuration c (tests of each configuration)
    if (const char* c = std::getenv("TTR_GEMM_CFG")) {
      const int v = std::atoi(c);
      TTR_REQUIRE(v >= 0 && v < 8, ST_INVALID_ARG, "TTR_GEMM_CFG must be 0..7");
      h->ctx.gemm_force_cfg = v;
    }
    // TTR_SVD_CERTIFY=0: run the SVD at every bond (no Cholesky certificate)
    if (const char* c = std::getenv("TTR_SVD_CERTIFY")) h->ctx.svd_certify = c[0] != '0';
    // TTR_JACOBI_TOL_SCALE=f (1e-3..1): scales the Jacobi rotation threshold max(p, 16) eps
    if (const char* c = std::getenv("TTR_JACOBI_TOL_SCALE")) {
      const double v = std::atof(c);
      TTR_REQUIRE(v >= 1e-3 && v <= 1.0, ST_INVALID_ARG, "TTR_JACOBI_TOL_SCALE must be in [1e-3, 1]");
      h->ctx.jacobi_tol_scale = v;
    }
    // TTR_LAUNCH_LOG=path: one line per GEMM dispatch (site, math precision, tile
    // configuration, workgroups, problems, algorithmic flops and bytes), appended to path
    if (const char* c = std::getenv("TTR_LAUNCH_LOG")) {
      h->ctx.launch_log = std::fopen(c, "a");
      TTR_REQUIRE(h->ctx.launch_log != nullptr, ST_INVALID_ARG, "TTR_LAUNCH_LOG: cannot open the file");
    }
    // TTR_GRAPHS=1: ttr_round_rand through HIP graphs (ttr_graph_enable)
    if (const char* c = std::getenv("TTR_GRAPHS")) h->graphs = c[0] == '1';
    // TTR_JACOBI_SORT=0: no column sorting between Jacobi sweeps
    if (const char* c = std::getenv("TTR_JACOBI_SORT")) h->ctx.jacobi_sort = c[0] != '0';
    // TTR_QR_PATH=p: 0 the QR path by size and batch, 1 the blocked path at every size, 2 the
    // blocked path with streaming panels, 3 the one-launch QR wherever it fits (tests of each path)
    if (const char* c = std::getenv("TTR_QR_PATH")) {
      const int v = std::atoi(c);
      TTR_REQUIRE(v >= 0 && v <= 3, ST_INVALID_ARG, "TTR_QR_PATH must be 0..3");
      h->ctx.qr_path = v;
    }
    // TTR_QR_GROUP=g (1, 2 or 4): 64-column blocks per compact WY group of the blocked QR
    if (const char* c = std::getenv("TTR_QR_GROUP")) {
      const int v = std::atoi(c);
      TTR_REQUIRE(v == 1 || v == 2 || v == 4, ST_INVALID_ARG, "TTR_QR_GROUP must be 1, 2 or 4");
      h->ctx.qr_group = v;
    }
    // TTR_JACOBI_MAX_SWEEPS=s (1..40): Jacobi sweeps before a non-converged SVD fails
    if (const char* c = std::getenv("TTR_JACOBI_MAX_SWEEPS")) {
      const int v = std::atoi(c);
      TTR_REQUIRE(v >= 1 && v <= 40, ST_INVALID_ARG, "TTR_JACOBI_MAX_SWEEPS must be 1..40");
      h->ctx.jacobi_max_sweeps = v;
    }
  } catch (const Error& e) {
    int st = e.status;
    if (h->ctx.stream) (void)hipStreamDestroy(h->ctx.stream);
    delete h;
    return st;
  }
  *out = h;
  return TTR_OK;
}

int ttr_destroy(ttr_handle_t h) {
  if (!h) return TTR_INVALID_ARG;
  (void)hipSetDevice(h->ctx.device);
  (void)hipStreamSynchronize(h->ctx.stream);
  h->graph_clear();
  for (auto& kv : h->tts) h->free_tt(kv.second.get());
  h->tts.clear();
  h->ops.clear();  // DevTTO frees its device copies
  (void)hipFree(h->one);
  if (h->ctx.launch_log) std::fclose(h->ctx.launch_log);
  if (h->stage) (void)hipHostFree(h->stage);
  if (h->ctx.splitk) (void)hipFree(h->ctx.splitk);
  h->ctx.ws.release();
  h->ctx.pin.release();
  h->ctx.cache.release_cached();
  (void)hipStreamDestroy(h->ctx.stream);
  delete h;
  return TTR_OK;
}

const char* ttr_last_error(ttr_handle_t h) { return h ? h->err.c_str() : "null handle"; }

int ttr_tt_upload(ttr_handle_t h, int64_t d, const int64_t* dims, const int64_t* rks, const int64_t* ot,
                  const double* const* cores, int64_t* id) {
  return guarded(h, [&] {
    TTR_REQUIRE(d >= 1 && dims && rks && cores && id, ST_INVALID_ARG, "bad arguments");
    std::vector<int64_t> vd(dims, dims + d), vr(rks, rks + d + 1), vo(size_t(d), 0);
    if (ot) vo.assign(ot, ot + d);
    for (int64_t k = 0; k < d; ++k) TTR_REQUIRE(vd[k] >= 1, ST_INVALID_ARG, "dims must be >= 1");
    for (int64_t k = 0; k <= d; ++k) TTR_REQUIRE(vr[k] >= 1, ST_INVALID_ARG, "ranks must be >= 1");
    TTR_REQUIRE(vr[0] == 1 && vr[d] == 1, ST_DIM_MISMATCH, "boundary ranks must be 1");
    auto t = h->make_tt(vd, vr, vo);
    int64_t total = 0;
    for (int k = 0; k < d; ++k) total += t->core_elems(k);
    double* stage = h->ctx.ws.alloc_n<double>(size_t(total));
    std::vector<Copy2D> cp;
    int64_t off = 0;
    for (int k = 0; k < d; ++k) {
      const int64_t n = vd[k], r = vr[k], r2 = vr[k + 1], e = t->core_elems(k);
      TTR_REQUIRE(cores[k] != nullptr, ST_INVALID_ARG, "null core pointer");
      TTR_HIP(hipMemcpy(stage + off, cores[k], size_t(e) * 8, hipMemcpyHostToDevice));
      Copy2D c = copy_job(stage + off, 1, n, t->cores[k], r, 1, int(n), int(r));
      c.ibs = n * r;
      c.obs = r * n;
      c.nb = int(r2);
      cp.push_back(c);
      off += e;
    }
    copy2d_batched(h->ctx, cp);
    *id = h->add(std::move(t));
  });
}

int ttr_tt_info(ttr_handle_t h, int64_t id, int64_t* d, int64_t* dims, int64_t* rks, int64_t* ot) {
  return guarded(h, [&] {
    DevTT* t = h->get(id);
    if (d) *d = t->d;
    if (dims) std::copy(t->dims.begin(), t->dims.end(), dims);
    if (rks) std::copy(t->rks.begin(), t->rks.end(), rks);
    if (ot) std::copy(t->ot.begin(), t->ot.end(), ot);
  });
}

int ttr_tt_download_core(ttr_handle_t h, int64_t id, int64_t k, double* dst) {
  return guarded(h, [&] {
    DevTT* t = h->get(id);
    TTR_REQUIRE(k >= 0 && k < t->d && dst, ST_INVALID_ARG, "bad core index");
    const int64_t n = t->dims[k], r = t->rks[k], r2 = t->rks[k + 1], e = t->core_elems(int(k));
    double* stage = h->ctx.ws.alloc_n<double>(size_t(std::max<int64_t>(e, 1)));
    Copy2D c = copy_job(t->cores[k], r, 1, stage, 1, n, int(n), int(r));
    c.ibs = r * n;
    c.obs = n * r;
    c.nb = int(r2);
    copy2d_batched(h->ctx, {c});
    h->ctx.sync();
    TTR_HIP(hipMemcpy(dst, stage, size_t(e) * 8, hipMemcpyDeviceToHost));
  });
}

int ttr_tt_free(ttr_handle_t h, int64_t id) {
  return guarded(h, [&] {
    h->ctx.sync();
    if (h->ops.count(id)) {  // operators share the id space
      h->ops.erase(id);
      return;
    }
    DevTT* t = h->get(id);
    h->free_tt(t);
    h->tts.erase(id);
  });
}

int ttr_tt_randn(ttr_handle_t h, int64_t d, const int64_t* dims, const int64_t* rks, uint64_t seed, int64_t* id) {
  return guarded(h, [&] {
    TTR_REQUIRE(d >= 1 && dims && rks && id, ST_INVALID_ARG, "bad arguments");
    std::vector<int64_t> vd(dims, dims + d), vr(rks, rks + d + 1);
    const int64_t rmax = *std::max_element(vr.begin(), vr.end());
    vr = r_and_d_to_rks(vr, vd, rmax);
    auto t = h->make_tt(vd, vr, std::vector<int64_t>(size_t(d), 0));
    for (int k = 0; k < d; ++k) randn_core(h->ctx, t->cores[k], int(vd[k]), int(vr[k]), int(vr[k + 1]), k, seed);
    *id = h->add(std::move(t));
  });
}

int ttr_tto_upload(ttr_handle_t h, int64_t d, const int64_t* dims, const int64_t* rks, const double* const* cores,
                   int64_t* id) {
  return guarded(h, [&] {
    TTR_REQUIRE(d >= 1 && dims && rks && cores && id, ST_INVALID_ARG, "bad arguments");
    auto op = std::make_unique<DevTTO>();
    op->d = int(d);
    op->dims.assign(dims, dims + d);
    op->rks.assign(rks, rks + d + 1);
    TTR_REQUIRE(op->rks[0] == 1 && op->rks[d] == 1, ST_DIM_MISMATCH, "boundary ranks must be 1");
    int64_t total = 0;
    for (int k = 0; k < d; ++k) total += op->dims[k] * op->dims[k] * op->rks[k] * op->rks[k + 1];
    TTR_HIP(hipMalloc(&op->base, size_t(std::max<int64_t>(total, 1)) * 8));
    int64_t off = 0;
    for (int k = 0; k < d; ++k) {
      const int64_t e = op->dims[k] * op->dims[k] * op->rks[k] * op->rks[k + 1];
      double* p = static_cast<double*>(op->base) + off;
      TTR_HIP(hipMemcpy(p, cores[k], size_t(e) * 8, hipMemcpyHostToDevice));
      op->cores.push_back(p);
      off += e;
    }
    // matrix copies of the wide cores, permuted on the host
    int64_t total_m = 0;
    for (int k = 0; k < d; ++k)
      if (op->wide(k)) total_m += op->dims[k] * op->dims[k] * op->rks[k] * op->rks[k + 1];
    op->hm.assign(size_t(d), nullptr);
    if (total_m > 0) {
      TTR_HIP(hipMalloc(&op->base_m, size_t(total_m) * 8));
      int64_t offm = 0;
      for (int k = 0; k < d; ++k) {
        if (!op->wide(k)) continue;
        const int64_t n = op->dims[k], R = op->rks[k], R1 = op->rks[k + 1];
        std::vector<double> m(size_t(n * n * R * R1));
        for (int64_t b = 0; b < R1; ++b)
          for (int64_t be = 0; be < R; ++be)
            for (int64_t j = 0; j < n; ++j)
              for (int64_t i = 0; i < n; ++i)
                m[size_t((i + n * b) + n * R1 * (be + R * j))] = cores[k][i + n * j + n * n * be + n * n * R * b];
        double* p = static_cast<double*>(op->base_m) + offm;
        TTR_HIP(hipMemcpy(p, m.data(), m.size() * 8, hipMemcpyHostToDevice));
        op->hm[size_t(k)] = p;
        offm += int64_t(m.size());
      }
    }
    const int64_t nid = h->next_id++;
    h->ops[nid] = std::move(op);
    *id = nid;
  });
}

int ttr_lincomb(ttr_handle_t h, int64_t nterms, const double* alpha, const int64_t* ids, int64_t* out_id) {
  return guarded(h, [&] {
    TTR_REQUIRE(nterms >= 1 && alpha && ids && out_id, ST_INVALID_ARG, "bad arguments");
    std::vector<DevTT*> xs;
    for (int64_t t = 0; t < nterms; ++t) {
      xs.push_back(h->get(ids[t]));
      check_dims_equal(xs[0], xs.back());
    }
    const int d = xs[0]->d;
    TTR_REQUIRE(d >= 2 || nterms == 1, ST_INVALID_ARG, "TT sums need d >= 2");
    // per term: ranks (a zero coefficient gives the rank-one zero TT, src/tt_operations.jl:432-433)
    std::vector<std::vector<int64_t>> tr(static_cast<size_t>(nterms));
    std::vector<int> scaled(static_cast<size_t>(nterms), 0);
    for (int64_t t = 0; t < nterms; ++t) {
      if (alpha[t] == 0.0) {
        tr[t].assign(size_t(d + 1), 1);
      } else {
        tr[t] = xs[t]->rks;
        auto it = std::find(xs[t]->ot.begin(), xs[t]->ot.end(), 0);
        TTR_REQUIRE(it != xs[t]->ot.end(), ST_INVALID_ARG, "no core with ot == 0 to scale");
        scaled[t] = int(it - xs[t]->ot.begin());
      }
    }
    std::vector<int64_t> rks(static_cast<size_t>(d + 1), 0);
    for (int k = 0; k <= d; ++k)
      for (int64_t t = 0; t < nterms; ++t) rks[k] += tr[t][k];
    rks[0] = 1;
    rks[d] = 1;
    if (nterms == 1) rks = tr[0];
    auto out = h->make_tt(xs[0]->dims, rks, std::vector<int64_t>(size_t(d), nterms == 1 ? 0 : 0));
    if (nterms == 1 && alpha[0] != 0.0) out->ot = xs[0]->ot;
    zeros_like_ranks(h, out.get());
    std::vector<BlockPut> bp;
    std::vector<int64_t> a0(static_cast<size_t>(d + 1), 0);
    for (int64_t t = 0; t < nterms; ++t) {
      if (alpha[t] != 0.0)
        for (int k = 0; k < d; ++k) {
          const int64_t ra = tr[t][k], rb = tr[t][k + 1];
          BlockPut b{};
          b.in = xs[t]->cores[k];
          b.out = out->cores[k];
          b.ra = int(ra);
          b.rb = int(rb);
          b.n = int(xs[0]->dims[k]);
          b.RA = int(rks[k]);
          b.RB = int(rks[k + 1]);
          b.a0 = k == 0 ? 0 : int(a0[k]);
          b.b0 = k == d - 1 ? 0 : int(a0[k + 1]);
          b.scale = (k == scaled[t]) ? alpha[t] : 1.0;
          bp.push_back(b);
        }
      for (int k = 0; k <= d; ++k) a0[k] += tr[t][k];
    }
    block_put_batched(h->ctx, bp);
    *out_id = h->add(std::move(out));
  });
}

int ttr_mpo_apply(ttr_handle_t h, int64_t op_id, int64_t x_id, int64_t* out_id) {
  return guarded(h, [&] {
    TTR_REQUIRE(out_id, ST_INVALID_ARG, "bad arguments");
    DevTTO* A = h->get_op(op_id);
    DevTT* x = h->get(x_id);
    TTR_REQUIRE(A->dims == x->dims, ST_DIM_MISMATCH, "Incompatible dimensions");
    std::vector<int64_t> rks(static_cast<size_t>(x->d + 1));
    for (int k = 0; k <= x->d; ++k) rks[k] = A->rks[k] * x->rks[k];
    auto y = h->make_tt(x->dims, rks, std::vector<int64_t>(size_t(x->d), 0));
    std::vector<MpoCore> jobs;
    for (int k = 0; k < x->d; ++k)
      jobs.push_back(MpoCore{A->cores[k], x->cores[k], y->cores[k], int(x->dims[k]), int(A->rks[k]),
                             int(A->rks[k + 1]), int(x->rks[k]), int(x->rks[k + 1])});
    mpo_apply_batched(h->ctx, jobs);
    *out_id = h->add(std::move(y));
  });
}

int ttr_hadamard(ttr_handle_t h, int64_t m, const int64_t* ids, int32_t first_fastest, int64_t* out_id) {
  return guarded(h, [&] {
    TTR_REQUIRE(m >= 1 && m <= KRON_MAX && ids && out_id, ST_INVALID_ARG, "bad arguments (1 <= m <= 8)");
    std::vector<DevTT*> xs;
    for (int64_t f = 0; f < m; ++f) {
      xs.push_back(h->get(ids[f]));
      check_dims_equal(xs[0], xs.back());
    }
    // factor order of the kernel: fastest first (src/tt_operations.jl:273-290 takes
    // kron(x, y) per slice, so there the last factor runs fastest)
    if (!first_fastest) std::reverse(xs.begin(), xs.end());
    const int d = xs[0]->d;
    std::vector<int64_t> rks(static_cast<size_t>(d + 1), 1);
    for (int k = 0; k <= d; ++k)
      for (auto* x : xs) {
        rks[k] *= x->rks[k];
        TTR_REQUIRE(rks[k] < (int64_t(1) << 31), ST_INVALID_ARG, "product rank too large");
      }
    for (int k = 0; k < d; ++k)
      TTR_REQUIRE(rks[k] * xs[0]->dims[k] * rks[k + 1] < (int64_t(1) << 31), ST_INVALID_ARG,
                  "product core exceeds 2^31 entries");
    auto y = h->make_tt(xs[0]->dims, rks, std::vector<int64_t>(size_t(d), 0));
    std::vector<KronCore> jobs;
    for (int k = 0; k < d; ++k) {
      KronCore J{};
      J.out = y->cores[k];
      J.m = int32_t(m);
      J.n = int32_t(xs[0]->dims[k]);
      J.RA = int32_t(rks[k]);
      J.RB = int32_t(rks[k + 1]);
      for (int64_t f = 0; f < m; ++f) {
        J.in[f] = xs[size_t(f)]->cores[k];
        J.ra[f] = int32_t(xs[size_t(f)]->rks[k]);
        J.rb[f] = int32_t(xs[size_t(f)]->rks[k + 1]);
      }
      jobs.push_back(J);
    }
    kron_cores_batched(h->ctx, jobs);
    *out_id = h->add(std::move(y));
  });
}

int ttr_orthogonalize(ttr_handle_t h, int64_t n, const int64_t* ids, int64_t root, int64_t* out_ids) {
  return guarded(h, [&] {
    TTR_REQUIRE(n >= 1 && ids && out_ids, ST_INVALID_ARG, "bad arguments");
    std::vector<DevTT*> xs;
    for (int64_t i = 0; i < n; ++i) xs.push_back(h->get(ids[i]));
    auto outs = orthogonalize_batch(h, xs, int(root));
    for (int64_t i = 0; i < n; ++i) out_ids[i] = h->add(std::move(outs[i]));
  });
}

int ttr_round_svd(ttr_handle_t h, int64_t n, const int64_t* ids, double tol, int64_t rmax, int direction,
                  int64_t* out_ids) {
  return guarded(h, [&] {
    TTR_REQUIRE(n >= 1 && ids && out_ids, ST_INVALID_ARG, "bad arguments");
    TTR_REQUIRE(direction == TTR_DIR_LEFT || direction == TTR_DIR_RIGHT, ST_INVALID_ARG,
                "Invalid direction. Use direction=:right or direction=:left");
    TTR_REQUIRE(rmax >= 1, ST_INVALID_ARG, "rmax must be >= 1");
    std::vector<DevTT*> xs;
    for (int64_t i = 0; i < n; ++i) xs.push_back(h->get(ids[i]));
    // orthogonalise the ones that need it (grouped by d for the root)
    std::vector<std::unique_ptr<DevTT>> tmp(static_cast<size_t>(n));
    std::vector<DevTT*> ys(static_cast<size_t>(n));
    std::map<int, std::vector<int64_t>> need;
    for (int64_t i = 0; i < n; ++i) {
      const bool ok = direction == TTR_DIR_LEFT ? is_leftorth(xs[i]) : is_rightorth(xs[i]);
      if (ok) ys[i] = xs[i];
      else need[xs[i]->d].push_back(i);
    }
    for (auto& kv : need) {
      std::vector<DevTT*> grp;
      for (int64_t i : kv.second) grp.push_back(xs[i]);
      ProfScope ps(h->ctx, "phase:orthogonalize");
      auto o = orthogonalize_batch(h, grp, direction == TTR_DIR_LEFT ? kv.first : 1);
      for (size_t g = 0; g < kv.second.size(); ++g) {
        ys[kv.second[g]] = o[g].get();
        tmp[kv.second[g]] = std::move(o[g]);
      }
    }
    std::vector<SweepState> st(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) st[i].y = ys[i];
    {
      ProfScope ps(h->ctx, "phase:svd_sweep");
      svd_sweep(h, st, tol, rmax, direction);
    }
    h->ctx.sync();
    for (auto& t : tmp)
      if (t) h->free_tt(t.get());
    for (int64_t i = 0; i < n; ++i) out_ids[i] = h->add(std::move(st[i].out));
  });
}

int ttr_norm(ttr_handle_t h, int64_t n, const int64_t* ids, double* out) {
  return guarded(h, [&] {
    TTR_REQUIRE(n >= 1 && ids && out, ST_INVALID_ARG, "bad arguments");
    std::vector<DevTT*> xs;
    for (int64_t i = 0; i < n; ++i) xs.push_back(h->get(ids[i]));
    std::vector<const double*> ptrs(static_cast<size_t>(n));
    std::vector<int64_t> cnt(static_cast<size_t>(n));
    std::vector<std::unique_ptr<DevTT>> tmp(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
      const DevTT* x = xs[i];
      int zeros = 0, root = -1;
      for (int k = 0; k < x->d; ++k)
        if (x->ot[k] == 0) { ++zeros; root = k; }
      if (zeros == 1) {
        ptrs[i] = x->cores[root];
        cnt[i] = x->core_elems(root);
      } else {
        auto o = orthogonalize_batch(h, {xs[i]}, x->d);
        tmp[i] = std::move(o[0]);
        ptrs[i] = tmp[i]->cores[x->d - 1];
        cnt[i] = tmp[i]->core_elems(x->d - 1);
      }
    }
    std::vector<double> r = frob_norms_host(h, ptrs, cnt);
    for (int64_t i = 0; i < n; ++i) out[i] = r[i];
    for (auto& t : tmp)
      if (t) h->free_tt(t.get());
  });
}

namespace {
// ------------------------------------------------------------ inner products

// Environments of pairs of TTs (x_i, y_i) with equal dims, all pairs in the same
// launches.  Left to right (dot, src/tt_operations.jl:302-321, and
// partial_contraction(; reverse=false), src/sketches.jl:1063-1067), with
// W_0 = 1 and W_k the rA_k x rB_k environment:
//   T       = W_k Y_k      (Y_k read as rB_k x (n rB_{k+1}))
//   W_{k+1} = X_k^T T      (X_k read as (rA_k n) x rA_{k+1}, T as (rA_k n) x rB_{k+1})
// Right to left (partial_contraction(; reverse=true), :1058-1062), W_d = 1:
//   T       = Y_k W_{k+1}^T   (Y_k read as (rB_k n) x rB_{k+1})
//   W_k     = X_k T^T         (X_k read as rA_k x (n rA_{k+1}), T as rB_k x (n rA_{k+1}))
// W[i][k] are device arrays (column-major) the caller provides for k = 0..d;
// the boundary environment (W_0 left to right, W_d right to left) is never read,
// the first core's product with it is skipped.
void pair_environments(ttr_handle_s* h, const std::vector<DevTT*>& xs, const std::vector<DevTT*>& ys, bool reverse,
                       const std::vector<std::vector<double*>>& W) {
  Ctx& ctx = h->ctx;
  const size_t np = xs.size();
  int dmax = 0;
  std::vector<double*> T(np);
  for (size_t i = 0; i < np; ++i) {
    const DevTT* x = xs[i];
    const DevTT* y = ys[i];
    int64_t mt = 1;
    for (int k = 0; k < x->d; ++k)
      mt = std::max(mt, reverse ? y->rks[k] * x->dims[k] * x->rks[k + 1] : x->rks[k] * x->dims[k] * y->rks[k + 1]);
    T[i] = ctx.ws.alloc_n<double>(size_t(mt));
    dmax = std::max(dmax, x->d);
  }
  GemmBatch gb;
  for (int s = 0; s < dmax; ++s) {
    // step s visits core k = s (left to right) or k = d-1-s (right to left)
    for (size_t i = 0; i < np; ++i) {
      const DevTT* x = xs[i];
      const DevTT* y = ys[i];
      if (s == 0 || s >= x->d) continue;
      const int k = reverse ? x->d - 1 - s : s;
      const int64_t n = x->dims[k], ra = x->rks[k], rb = y->rks[k], ra1 = x->rks[k + 1], rb1 = y->rks[k + 1];
      if (reverse)
        gb.add(int(rb * n), int(ra1), int(rb1), 1.0, y->cores[k], 1, rb * n, W[i][size_t(k + 1)], ra1, 1, 0.0, T[i],
               1, rb * n);
      else
        gb.add(int(ra), int(n * rb1), int(rb), 1.0, W[i][size_t(k)], 1, ra, y->cores[k], 1, rb, 0.0, T[i], 1, ra);
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "dot_env");
    for (size_t i = 0; i < np; ++i) {
      const DevTT* x = xs[i];
      const DevTT* y = ys[i];
      if (s >= x->d) continue;
      const int k = reverse ? x->d - 1 - s : s;
      const int64_t n = x->dims[k], ra = x->rks[k], rb = y->rks[k], ra1 = x->rks[k + 1], rb1 = y->rks[k + 1];
      // the boundary environment is 1: the first product's result is Y_k itself
      const double* Tk = s == 0 ? y->cores[k] : T[i];
      if (reverse)
        gb.add(int(ra), int(rb), int(n * ra1), 1.0, x->cores[k], 1, ra, Tk, rb, 1, 0.0, W[i][size_t(k)], 1, ra);
      else
        gb.add(int(ra1), int(rb1), int(ra * n), 1.0, x->cores[k], ra * n, 1, Tk, 1, ra * n, 0.0, W[i][size_t(k + 1)],
               1, ra1);
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "dot_core");
  }
}

// <psi_i|H_i|phi_i> (dot_operator, src/tt_operations.jl:385-421) for a batch of
// triples, folded left to right through the three-index environment
// L[(a, b, c)] (psi bond fastest, then the H bond, then the phi bond), all
// triples in the same launches:
//   T1 = L B_k                         GEMM, L read as (rA R) x rB, B_k as rB x (n rB')
//   T2[(a, i), (b', c)] = sum_{b, j} H_k[i, j, b, b'] T1[(a, b), (j, c)]
//                                      op_apply, or for wide bonds one GEMM per c
//                                      against the matrix copy Hm of H_k
//   L' = A_k^T T2                      GEMM, A_k read as (rA n) x rA'
// The reference pivots at the orthogonality centre and folds from both ends;
// the contraction is the same sum, so results agree to rounding.
void dot_operator_batch(ttr_handle_s* h, const std::vector<DevTT*>& ps, const std::vector<DevTTO*>& Hs,
                        const std::vector<DevTT*>& fs, double* out) {
  Ctx& ctx = h->ctx;
  const size_t nt = ps.size();
  int dmax = 0;
  std::vector<double*> Lb(nt), T1(nt), T2(nt);
  for (size_t i = 0; i < nt; ++i) {
    const DevTT* a = ps[i];
    const DevTTO* H = Hs[i];
    const DevTT* b = fs[i];
    int64_t ml = 1, m1 = 1, m2 = 1;
    for (int k = 0; k < a->d; ++k) {
      const int64_t n = a->dims[k];
      ml = std::max(ml, a->rks[k + 1] * H->rks[k + 1] * b->rks[k + 1]);
      m1 = std::max(m1, a->rks[k] * H->rks[k] * n * b->rks[k + 1]);
      m2 = std::max(m2, a->rks[k] * n * H->rks[k + 1] * b->rks[k + 1]);
    }
    Lb[i] = ctx.ws.alloc_n<double>(size_t(ml));
    T1[i] = ctx.ws.alloc_n<double>(size_t(m1));
    T2[i] = ctx.ws.alloc_n<double>(size_t(m2));
    dmax = std::max(dmax, a->d);
  }
  GemmBatch gb;
  std::vector<OpApply> oj;
  for (int k = 0; k < dmax; ++k) {
    if (k > 0) {
      for (size_t i = 0; i < nt; ++i) {
        const DevTT* a = ps[i];
        const DevTT* b = fs[i];
        if (k >= a->d) continue;
        const int64_t n = a->dims[k], rar = a->rks[k] * Hs[i]->rks[k], rb = b->rks[k], rb1 = b->rks[k + 1];
        gb.add(int(rar), int(n * rb1), int(rb), 1.0, Lb[i], 1, rar, b->cores[k], 1, rb, 0.0, T1[i], 1, rar);
      }
      gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "dotop_l");
    }
    for (size_t i = 0; i < nt; ++i) {
      const DevTT* a = ps[i];
      const DevTTO* H = Hs[i];
      const DevTT* b = fs[i];
      if (k >= a->d) continue;
      const int64_t n = a->dims[k], ra = a->rks[k], R = H->rks[k], R1 = H->rks[k + 1], rb1 = b->rks[k + 1];
      const double* in = k == 0 ? b->cores[0] : T1[i];
      if (!H->wide(k)) {
        OpApply o = op_job(H, k, false);
        o.in = in;
        o.out = T2[i];
        o.X = int(ra);
        o.Y = int(rb1);
        o.is_x = 1; o.is_b = ra; o.is_m = ra * R; o.is_y = ra * R * n;
        o.os_x = 1; o.os_m = ra; o.os_b = ra * n; o.os_y = ra * n * R1;
        oj.push_back(o);
      } else {
        for (int64_t c = 0; c < rb1; ++c)
          gb.add(int(ra), int(n * R1), int(R * n), 1.0, in + ra * R * n * c, 1, ra, H->hm[size_t(k)], n * R1, 1, 0.0,
                 T2[i] + ra * n * R1 * c, 1, ra);
      }
    }
    op_apply_batched(ctx, oj, Prec::F64, Prec::F64);
    oj.clear();
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "dotop_h");
    for (size_t i = 0; i < nt; ++i) {
      const DevTT* a = ps[i];
      const DevTT* b = fs[i];
      if (k >= a->d) continue;
      const int64_t n = a->dims[k], ra = a->rks[k], ra1 = a->rks[k + 1], cols = Hs[i]->rks[k + 1] * b->rks[k + 1];
      double* C = k == a->d - 1 ? out + i : Lb[i];
      gb.add(int(ra1), int(cols), int(ra * n), 1.0, a->cores[k], ra * n, 1, T2[i], 1, ra * n, 0.0, C, 1, ra1);
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "dotop_r");
  }
}
}  // namespace

int ttr_dot(ttr_handle_t h, int64_t n, const int64_t* x_ids, const int64_t* y_ids, double* out) {
  return guarded(h, [&] {
    TTR_REQUIRE(n >= 1 && x_ids && y_ids && out, ST_INVALID_ARG, "bad arguments");
    std::vector<DevTT*> xs, ys;
    std::vector<std::vector<double*>> W(static_cast<size_t>(n));
    double* res = h->ctx.ws.alloc_n<double>(size_t(n));
    for (int64_t i = 0; i < n; ++i) {
      xs.push_back(h->get(x_ids[i]));
      ys.push_back(h->get(y_ids[i]));
      TTR_REQUIRE(xs.back()->dims == ys.back()->dims, ST_DIM_MISMATCH, "TT dimensions are not compatible");
      const DevTT* x = xs.back();
      const DevTT* y = ys.back();
      W[i].assign(size_t(x->d + 1), nullptr);
      for (int k = 1; k < x->d; ++k) W[i][size_t(k)] = h->ctx.ws.alloc_n<double>(size_t(x->rks[k] * y->rks[k]));
      W[i][size_t(x->d)] = res + i;
    }
    pair_environments(h, xs, ys, false, W);
    h->ctx.sync();
    TTR_HIP(hipMemcpy(out, res, size_t(n) * 8, hipMemcpyDeviceToHost));
  });
}

int ttr_dot_noalloc(ttr_handle_t h, int64_t n, const int64_t* x_ids, const int64_t* y_ids, double* out) {
  // the workspace and staging arenas point at the scratch array and the handle's
  // pinned staging for this call only
  struct FixedScope {
    Ctx& c;
    FixedScope(Ctx& ctx, void* dev, size_t dcap, void* host, size_t hcap) : c(ctx) {
      c.ws.set_fixed(dev, dcap);
      c.pin.set_fixed(host, hcap);
      c.fsplit = nullptr;
      c.fsplit_elems = 0;
    }
    ~FixedScope() {
      (void)hipStreamSynchronize(c.stream);  // nothing may still read the scratch
      c.ws.clear_fixed();
      c.pin.clear_fixed();
      c.fsplit = nullptr;
      c.fsplit_elems = 0;
    }
  };
  return guarded(h, [&] {
    TTR_REQUIRE(n >= 1 && x_ids && y_ids && out, ST_INVALID_ARG, "bad arguments");
    TTR_REQUIRE(h->scratch && h->stage, ST_DEVICE_ERROR, "the handle has no scratch");
    std::vector<DevTT*> xs, ys;
    // environments: two per pair (W_k and W_{k+1}, alternating), the product
    // temporary and the result; 256-byte granules as the arena hands them out
    auto gran = [](int64_t elems) { return (size_t(std::max<int64_t>(elems, 1)) * 8 + 255) & ~size_t(255); };
    size_t need = gran(n);
    for (int64_t i = 0; i < n; ++i) {
      xs.push_back(h->get(x_ids[i]));
      ys.push_back(h->get(y_ids[i]));
      TTR_REQUIRE(xs.back()->dims == ys.back()->dims, ST_DIM_MISMATCH, "TT dimensions are not compatible");
      const DevTT* x = xs.back();
      const DevTT* y = ys.back();
      int64_t mw = 1, mt = 1;
      for (int k = 0; k < x->d; ++k) {
        mw = std::max(mw, x->rks[k] * y->rks[k]);
        mt = std::max(mt, x->rks[k] * x->dims[k] * y->rks[k + 1]);
      }
      need += 2 * gran(mw) + gran(mt);
    }
    TTR_REQUIRE(need <= TTR_SCRATCH_BYTES / 2, ST_INVALID_ARG,
                "ttr_dot_noalloc: the environments need " + std::to_string(need >> 20) + " MiB; the limit is " +
                    std::to_string(TTR_SCRATCH_BYTES >> 21) + " MiB of the " + std::to_string(TTR_SCRATCH_BYTES >> 20) +
                    " MiB device scratch (the rest holds descriptors and split-K sums)");
    FixedScope fs(h->ctx, h->scratch, TTR_SCRATCH_BYTES, h->stage, TTR_STAGE_BYTES);
    double* res = h->ctx.ws.alloc_n<double>(size_t(n));
    std::vector<std::vector<double*>> W(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
      const DevTT* x = xs[i];
      const DevTT* y = ys[i];
      int64_t mw = 1;
      for (int k = 0; k < x->d; ++k) mw = std::max(mw, x->rks[k] * y->rks[k]);
      double* pp[2] = {h->ctx.ws.alloc_n<double>(size_t(mw)), h->ctx.ws.alloc_n<double>(size_t(mw))};
      W[i].assign(size_t(x->d + 1), nullptr);
      for (int k = 1; k < x->d; ++k) W[i][size_t(k)] = pp[k & 1];  // step k reads W_k, writes W_{k+1}
      W[i][size_t(x->d)] = res + i;
    }
    pair_environments(h, xs, ys, false, W);
    h->ctx.sync();
    TTR_HIP(hipMemcpy(out, res, size_t(n) * 8, hipMemcpyDeviceToHost));
  });
}

int ttr_partial_contraction(ttr_handle_t h, int64_t x_id, int64_t y_id, int reverse, double* const* W_out) {
  return guarded(h, [&] {
    TTR_REQUIRE(W_out, ST_INVALID_ARG, "bad arguments");
    DevTT* x = h->get(x_id);
    DevTT* y = h->get(y_id);
    TTR_REQUIRE(x->dims == y->dims, ST_DIM_MISMATCH, "TT dimensions are not compatible");
    const int d = x->d;
    std::vector<std::vector<double*>> W(1, std::vector<double*>(size_t(d + 1)));
    for (int k = 0; k <= d; ++k) W[0][size_t(k)] = h->ctx.ws.alloc_n<double>(size_t(x->rks[k] * y->rks[k]));
    pair_environments(h, {x}, {y}, reverse != 0, W);
    h->ctx.sync();
    for (int k = 0; k <= d; ++k) {
      if (!W_out[k]) continue;
      const bool boundary = reverse ? k == d : k == 0;
      if (boundary) W_out[k][0] = 1.0;
      else TTR_HIP(hipMemcpy(W_out[k], W[0][size_t(k)], size_t(x->rks[k] * y->rks[k]) * 8, hipMemcpyDeviceToHost));
    }
  });
}

int ttr_dot_operator(ttr_handle_t h, int64_t n, const int64_t* psi_ids, const int64_t* op_ids, const int64_t* phi_ids,
                     double* out) {
  return guarded(h, [&] {
    TTR_REQUIRE(n >= 1 && psi_ids && op_ids && phi_ids && out, ST_INVALID_ARG, "bad arguments");
    std::vector<DevTT*> ps, fs;
    std::vector<DevTTO*> Hs;
    for (int64_t i = 0; i < n; ++i) {
      ps.push_back(h->get(psi_ids[i]));
      Hs.push_back(h->get_op(op_ids[i]));
      fs.push_back(h->get(phi_ids[i]));
      TTR_REQUIRE(ps.back()->dims == fs.back()->dims && Hs.back()->dims == fs.back()->dims, ST_DIM_MISMATCH,
                  "TT dimensions are not compatible");
    }
    double* res = h->ctx.ws.alloc_n<double>(size_t(n));
    dot_operator_batch(h, ps, Hs, fs, res);
    h->ctx.sync();
    TTR_HIP(hipMemcpy(out, res, size_t(n) * 8, hipMemcpyDeviceToHost));
  });
}

namespace {
// ------------------------------------------------- entrywise (Hadamard) products
// One NTuple item: the entrywise product x_1 .* ... .* x_M of TTs of equal dims,
// never formed.  Its bond k carries the factors' rank indices with the first
// factor fastest (P_k = prod_m r_mk), the order of the reference's NTuple
// methods (src/tt_randtools.jl:378-382, src/sketches.jl:574-575).
struct TupleItem {
  std::vector<DevTT*> f;
  std::vector<int64_t> rks, p, s;  // target sketch ranks, block counts, sketch ranks (d+1)
  std::vector<void*> W;            // W_k: P_k x s_k in the working precision
  int32_t* E = nullptr;            // per core exponents of W (d+1)
  int64_t P(int k) const {
    int64_t v = 1;
    for (auto* x : f) v *= x->rks[k];
    return v;
  }
};

// T1[(a_1..a_M), i, c] = sum_al prod_m A_m[a_m, i, al_m] V[(al_1..al_M), c] for
// every item (V: P_{k+1} x ncol, working precision; T1: P_k x n x ncol), one
// factor at a time (contract_sketch_core_kronecker_backwards!, src/sketches.jl:
// 1735-1881, whose cost is n * sum_m (prod of ranks) instead of the product core's
// n * P_k * P_{k+1}): factor m contracts its right rank index for every mode
// index i as one problem per (i, remaining column) of a grouped GEMM.
// Intermediates alternate between U[0] and U[1] (item-sized regions).
void tuple_apply_cores_right(Ctx& ctx, std::vector<TupleItem>& its, int k, const std::vector<const void*>& V,
                             const std::vector<int64_t>& ncol, const std::vector<void*>& T1,
                             const std::vector<std::array<void*, 2>>& U, Prec wp, const char* site) {
  const size_t eb = prec_bytes(wp);
  size_t maxM = 0;
  for (auto& it : its) maxM = std::max(maxM, it.f.size());
  GemmBatch gb;
  for (size_t m = 0; m < maxM; ++m) {
    for (size_t q = 0; q < its.size(); ++q) {
      TupleItem& it = its[q];
      const size_t M = it.f.size();
      if (m >= M) continue;
      const int64_t n = it.f[0]->dims[k];
      const DevTT* x = it.f[m];
      const int64_t L = x->rks[k], R = x->rks[k + 1];
      int64_t pre = 1, rest = ncol[q];
      for (size_t j = 0; j < m; ++j) pre *= it.f[j]->rks[k];
      for (size_t j = m + 1; j < M; ++j) rest *= it.f[j]->rks[k + 1];
      const int64_t Pk = it.P(k);
      const bool last = m + 1 == M;
      const char* src = static_cast<const char*>(m == 0 ? V[q] : U[q][(m - 1) & 1]);
      char* dst = last ? static_cast<char*>(T1[q]) : static_cast<char*>(U[q][m & 1]);
      for (int64_t i = 0; i < n; ++i) {
        const double* Ai = x->cores[k] + L * i;  // A(a, al) = core[a + L (i + n al)]
        if (m == 0) {
          // U_1[i] (L x rest) = A_i (L x R) V (R x rest)
          if (last) gb.add(int(L), int(rest), int(R), 1.0, Ai, 1, L * n, src, 1, R, 0.0, dst + size_t(L * i) * eb, 1,
                           L * n);
          else gb.add(int(L), int(rest), int(R), 1.0, Ai, 1, L * n, src, 1, R, 0.0,
                      dst + size_t(i * L * rest) * eb, 1, L);
          continue;
        }
        // per remaining column r: C^T (L x pre) = A_i (L x R) U_slab^T (R x pre)
        const char* Ui = src + size_t(i * pre * R * rest) * eb;
        for (int64_t r = 0; r < rest; ++r) {
          const char* slab = Ui + size_t(r * pre * R) * eb;
          if (last) gb.add(int(L), int(pre), int(R), 1.0, Ai, 1, L * n, slab, pre, 1, 0.0,
                           dst + size_t(Pk * (i + n * r)) * eb, pre, 1);
          else gb.add(int(L), int(pre), int(R), 1.0, Ai, 1, L * n, slab, pre, 1, 0.0,
                      dst + size_t(i * pre * L * rest + r * pre * L) * eb, pre, 1);
        }
      }
    }
    gemm_run(ctx, gb, Prec::F64, wp, wp, wp, site);
  }
}

// Right-to-left recursive sketch of every item (tt_recursive_sketch(T,
// A::NTuple, rks; reverse=true), src/sketches.jl:474-535): W_k from W_{k+1} with
// the factored core contraction, then the block products with S_k as in the
// plain sketch; W_k is normalised to powers of two per core (exponents in E).
void tuple_sketch_phase(ttr_handle_s* h, SketchPlan& sp, std::vector<TupleItem>& its, int kmin) {
  Ctx& ctx = h->ctx;
  const int d = sp.d;
  const Prec wp = sp.wprec;
  const size_t eb = prec_bytes(wp), es = prec_bytes(sp.prec);
  std::vector<void*> T1(its.size());
  std::vector<std::array<void*, 2>> U(its.size());
  std::vector<double*> last(its.size());
  int32_t* E_all = ctx.ws.alloc_n<int32_t>(its.size() * size_t(d + 1));
  zero_fill(ctx, E_all, its.size() * size_t(d + 1) * 4);
  for (size_t q = 0; q < its.size(); ++q) {
    TupleItem& it = its[q];
    it.W.assign(size_t(d + 1), nullptr);
    it.E = E_all + q * size_t(d + 1);
    int64_t t1 = 1, u = 1;
    for (int k = kmin; k < d; ++k) {
      it.W[size_t(k)] = ctx.ws.alloc(size_t(it.P(k) * it.s[size_t(k)]) * eb);
      if (k + 1 < d) {
        const int64_t n = sp.dims[k], cols = it.s[size_t(k + 1)];
        t1 = std::max(t1, it.P(k) * n * cols);
        // intermediate after factor m: n * (prod_{j<=m} L_j) * (prod_{j>m} R_j) * cols
        for (size_t m = 0; m + 1 < it.f.size(); ++m) {
          int64_t v = n * cols;
          for (size_t j = 0; j <= m; ++j) v *= it.f[j]->rks[k];
          for (size_t j = m + 1; j < it.f.size(); ++j) v *= it.f[j]->rks[k + 1];
          u = std::max(u, v);
        }
      }
    }
    T1[q] = ctx.ws.alloc(size_t(t1) * eb);
    U[q] = {ctx.ws.alloc(size_t(u) * eb), ctx.ws.alloc(size_t(u) * eb)};
    last[q] = ctx.ws.alloc_n<double>(size_t(it.P(d - 1) * sp.dims[d - 1]));
  }
  // the last core of each product, formed explicitly (P_{d-1} x n x 1)
  std::vector<KronCore> kc;
  for (size_t q = 0; q < its.size(); ++q) {
    KronCore J{};
    J.out = last[q];
    J.m = int32_t(its[q].f.size());
    J.n = int32_t(sp.dims[d - 1]);
    J.RA = int32_t(its[q].P(d - 1));
    J.RB = 1;
    for (size_t f = 0; f < its[q].f.size(); ++f) {
      J.in[f] = its[q].f[f]->cores[d - 1];
      J.ra[f] = int32_t(its[q].f[f]->rks[d - 1]);
      J.rb[f] = 1;
    }
    kc.push_back(J);
  }
  kron_cores_batched(ctx, kc);
  GemmBatch gb;
  for (int k = d - 1; k >= kmin; --k) {
    const int64_t n = sp.dims[k], bvr = sp.bv[k + 1], bvl = sp.bv[k];
    const char* S = static_cast<const char*>(sp.S[k]);
    if (k == d - 1) {
      for (size_t q = 0; q < its.size(); ++q) {
        const int64_t Pk = its[q].P(k);
        for (int64_t j = 0; j < its[q].p[size_t(k)]; ++j)
          gb.add(int(Pk), int(bvl), int(n), 1.0, last[q], 1, Pk, S + size_t(n * bvl * j) * es, 1, n, 0.0,
                 static_cast<char*>(its[q].W[size_t(k)]) + size_t(Pk * bvl * j) * eb, 1, Pk);
      }
      gemm_run(ctx, gb, Prec::F64, sp.prec, wp, wp, "tuple_sketch_last");
    } else {
      std::vector<const void*> V;
      std::vector<int64_t> cols;
      for (auto& it : its) {
        V.push_back(it.W[size_t(k + 1)]);
        cols.push_back(it.s[size_t(k + 1)]);
      }
      tuple_apply_cores_right(ctx, its, k, V, cols, T1, U, wp, "tuple_sketch_cores");
      for (size_t q = 0; q < its.size(); ++q) {
        const int64_t Pk = its[q].P(k);
        const char* A = static_cast<const char*>(T1[q]);
        for (int64_t j = 0; j < its[q].p[size_t(k)]; ++j)
          gb.add(int(Pk), int(bvl), int(n * bvr), 1.0, A + size_t(Pk * n * bvr * j) * eb, 1, Pk,
                 S + size_t(n * bvr * bvl * j) * es, 1, n * bvr, 0.0,
                 static_cast<char*>(its[q].W[size_t(k)]) + size_t(Pk * bvl * j) * eb, 1, Pk);
      }
      gemm_run(ctx, gb, wp, sp.prec, wp, wp, "tuple_sketch_w");
    }
    std::vector<Pow2Job> pj;
    for (auto& it : its)
      pj.push_back(Pow2Job{it.W[size_t(k)], it.P(k), int(it.P(k)), int(it.s[size_t(k)]),
                           k == d - 1 ? nullptr : it.E + k + 1, it.E + k});
    pow2_normalize_batched(ctx, pj, wp);
  }
}

// Sketch plan and items of a batch of NTuple items (sketch structure from the
// target ranks, as for plain TTs)
std::vector<TupleItem> tuple_items(ttr_handle_s* h, int64_t n, const int64_t* nfactors, const int64_t* ids,
                                   const int64_t* rks, const ttr_rand_opts* o, SketchPlan& sp) {
  TTR_REQUIRE(n >= 1 && nfactors && ids && rks, ST_INVALID_ARG, "bad arguments");
  std::vector<TupleItem> its(static_cast<size_t>(n));
  std::vector<RandItem> plan(static_cast<size_t>(n));
  int64_t off = 0;
  const DevTT* x0 = nullptr;
  for (int64_t i = 0; i < n; ++i) {
    TTR_REQUIRE(nfactors[i] >= 1 && nfactors[i] <= KRON_MAX, ST_INVALID_ARG, "an item needs 1..8 factors");
    for (int64_t f = 0; f < nfactors[i]; ++f) {
      DevTT* x = h->get(ids[off++]);
      if (!x0) x0 = x;
      check_dims_equal(x0, x);
      TTR_REQUIRE(x->rks[0] == 1 && x->rks[size_t(x->d)] == 1, ST_INVALID_ARG, "factors need rks[0] = rks[d] = 1");
      its[size_t(i)].f.push_back(x);
    }
  }
  const int d = x0->d;
  for (int64_t i = 0; i < n; ++i) {
    plan[size_t(i)].rks.assign(rks + i * (d + 1), rks + (i + 1) * (d + 1));
    for (int k = 0; k <= d; ++k)
      TTR_REQUIRE(its[size_t(i)].P(k) < (int64_t(1) << 31), ST_INVALID_ARG, "product rank too large");
  }
  prepare_plan(h, o, d, x0->dims, plan, sp);
  for (int64_t i = 0; i < n; ++i) {
    its[size_t(i)].rks = plan[size_t(i)].rks;
    its[size_t(i)].p = plan[size_t(i)].p;
    its[size_t(i)].s = plan[size_t(i)].s;
  }
  return its;
}

// ttrand_rounding(y::NTuple{M,TTvector}, rks) (src/tt_randtools.jl:364-458),
// batched: the factored right-to-left sketch, then per core k (FP64)
//   Z_k = y_k W_{k+1}, Q_k = qr(Z_k), Qy = Q_k^T y_k,
//   y_{k+1}[(rho, i), (w_1..w_M)] = sum_l Qy[rho, (l_1..l_M)] prod_m A_m[l_m, i, w_m]
// where the left part y_k ((r_k n_k) x P_{k+1}) carries the factors' rank
// indices and the next cores are applied one factor at a time.
std::vector<std::unique_ptr<DevTT>> round_rand_tuple_batch(ttr_handle_s* h, std::vector<TupleItem>& its,
                                                           SketchPlan& sp) {
  Ctx& ctx = h->ctx;
  const int d = sp.d;
  std::vector<std::unique_ptr<DevTT>> outs(its.size());
  std::vector<std::vector<int64_t>> orr(its.size());
  for (size_t q = 0; q < its.size(); ++q) {
    std::vector<int64_t> r(size_t(d + 1), 1);
    for (int k = 0; k + 1 < d; ++k) r[size_t(k + 1)] = std::min(sp.dims[k] * r[size_t(k)], its[q].s[size_t(k + 1)]);
    std::vector<int64_t> ot(size_t(d), 1);
    ot[size_t(d - 1)] = 0;
    outs[q] = h->make_tt(sp.dims, r, ot);
    orr[q] = r;
  }
  if (d == 1) {  // the product itself (one core)
    std::vector<KronCore> kc;
    for (size_t q = 0; q < its.size(); ++q) {
      KronCore J{};
      J.out = outs[q]->cores[0];
      J.m = int32_t(its[q].f.size());
      J.n = int32_t(sp.dims[0]);
      J.RA = J.RB = 1;
      for (size_t f = 0; f < its[q].f.size(); ++f) {
        J.in[f] = its[q].f[f]->cores[0];
        J.ra[f] = J.rb[f] = 1;
      }
      kc.push_back(J);
    }
    kron_cores_batched(ctx, kc);
    return outs;
  }
  {
    ProfScope ps(ctx, "phase:reverse_sketch");
    make_sketch(h, sp, 1);
    tuple_sketch_phase(h, sp, its, 1);
  }
  ProfScope ps_orth(ctx, "phase:orthogonalization");
  struct Buf {
    double* y[2];
    double* Z;
    double* Qy;
    double* V[2];
  };
  std::vector<Buf> bf(its.size());
  for (size_t q = 0; q < its.size(); ++q) {
    const TupleItem& it = its[q];
    const size_t M = it.f.size();
    int64_t ymax = sp.dims[0] * it.P(1), zmax = 1, qmax = 1, vmax = 1;
    for (int k = 0; k + 1 < d; ++k) {
      const int64_t r = orr[q][size_t(k)], r1 = orr[q][size_t(k + 1)], n = sp.dims[k], P1 = it.P(k + 1);
      ymax = std::max(ymax, r * n * P1);
      zmax = std::max(zmax, r * n * it.s[size_t(k + 1)]);
      qmax = std::max(qmax, r1 * P1);
      const int kk = k + 1, n1 = int(sp.dims[kk]);
      for (size_t m = 0; m + 1 < M; ++m) {  // after factor m at core kk
        int64_t v = int64_t(n1) * r1;
        for (size_t j = 0; j <= m; ++j) v *= it.f[j]->rks[kk + 1];
        for (size_t j = m + 1; j < M; ++j) v *= it.f[j]->rks[kk];
        vmax = std::max(vmax, v);
      }
    }
    bf[q].y[0] = ctx.ws.alloc_n<double>(size_t(ymax));
    bf[q].y[1] = ctx.ws.alloc_n<double>(size_t(ymax));
    bf[q].Z = ctx.ws.alloc_n<double>(size_t(zmax));
    bf[q].Qy = ctx.ws.alloc_n<double>(size_t(qmax));
    bf[q].V[0] = ctx.ws.alloc_n<double>(size_t(vmax));
    bf[q].V[1] = ctx.ws.alloc_n<double>(size_t(vmax));
  }
  // y_0 = the first core of each product (1 x n_0 x P_1)
  {
    std::vector<KronCore> kc;
    for (size_t q = 0; q < its.size(); ++q) {
      KronCore J{};
      J.out = bf[q].y[0];
      J.m = int32_t(its[q].f.size());
      J.n = int32_t(sp.dims[0]);
      J.RA = 1;
      J.RB = int32_t(its[q].P(1));
      for (size_t f = 0; f < its[q].f.size(); ++f) {
        J.in[f] = its[q].f[f]->cores[0];
        J.ra[f] = 1;
        J.rb[f] = int32_t(its[q].f[f]->rks[1]);
      }
      kc.push_back(J);
    }
    kron_cores_batched(ctx, kc);
  }
  GemmBatch gb;
  int yb = 0;
  for (int k = 0; k + 1 < d; ++k) {
    const int64_t n = sp.dims[k];
    for (size_t q = 0; q < its.size(); ++q) {
      const int64_t r = orr[q][size_t(k)], s = its[q].s[size_t(k + 1)], P1 = its[q].P(k + 1);
      gb.add(int(r * n), int(s), int(P1), 1.0, bf[q].y[yb], 1, r * n, its[q].W[size_t(k + 1)], 1, P1, 0.0, bf[q].Z,
             1, r * n);
    }
    gemm_run(ctx, gb, Prec::F64, sp.wprec, Prec::F64, Prec::F64, "tuple_rr_Z");
    std::vector<QRJob> qj;
    for (size_t q = 0; q < its.size(); ++q) {
      const int64_t r = orr[q][size_t(k)], s = its[q].s[size_t(k + 1)];
      qj.push_back(QRJob{bf[q].Z, r * n, int(r * n), int(s), outs[q]->cores[k], r * n});
    }
    qr_batched(ctx, qj);
    // Qy = Q^T y_k (r_{k+1} x P_{k+1})
    for (size_t q = 0; q < its.size(); ++q) {
      const int64_t r = orr[q][size_t(k)], r1 = orr[q][size_t(k + 1)], P1 = its[q].P(k + 1);
      gb.add(int(r1), int(P1), int(r * n), 1.0, outs[q]->cores[k], r * n, 1, bf[q].y[yb], 1, r * n, 0.0, bf[q].Qy, 1,
             r1);
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "tuple_rr_P");
    // y_{k+1} (or the last output core): the factors' cores at k+1 applied to Qy
    const int kk = k + 1;
    const int64_t n1 = sp.dims[kk];
    size_t maxM = 0;
    for (auto& it : its) maxM = std::max(maxM, it.f.size());
    for (size_t m = 0; m < maxM; ++m) {
      for (size_t q = 0; q < its.size(); ++q) {
        const TupleItem& it = its[q];
        const size_t M = it.f.size();
        if (m >= M) continue;
        const int64_t r1 = orr[q][size_t(kk)];
        const DevTT* x = it.f[m];
        const int64_t L = x->rks[kk], Wm = x->rks[kk + 1];
        int64_t pre = r1, rest = 1;
        for (size_t j = 0; j < m; ++j) pre *= it.f[j]->rks[kk + 1];
        for (size_t j = m + 1; j < M; ++j) rest *= it.f[j]->rks[kk];
        const bool lastf = m + 1 == M;
        const double* src = m == 0 ? bf[q].Qy : bf[q].V[(m - 1) & 1];
        double* dst = lastf ? (kk == d - 1 ? outs[q]->cores[kk] : bf[q].y[yb ^ 1]) : bf[q].V[m & 1];
        for (int64_t i = 0; i < n1; ++i) {
          const double* Bi = x->cores[kk] + L * i;  // B(l, w) = core[l + L (i + n1 w)]
          const double* Ai = m == 0 ? src : src + i * pre * L * rest;
          if (!lastf) {
            for (int64_t c = 0; c < rest; ++c)
              gb.add(int(pre), int(Wm), int(L), 1.0, Ai + c * pre * L, 1, pre, Bi, 1, L * n1, 0.0,
                     dst + i * pre * Wm * rest + c * pre * Wm, 1, pre);
          } else {
            // rows (rho, w_1..w_{M-1}) split by the prefix t, columns w_M
            const int64_t npre = pre / r1;
            for (int64_t t = 0; t < npre; ++t)
              gb.add(int(r1), int(Wm), int(L), 1.0, Ai + t * r1, 1, pre, Bi, 1, L * n1, 0.0,
                     dst + r1 * i + r1 * n1 * t, 1, r1 * n1 * npre);
          }
        }
      }
      gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "tuple_rr_Y");
    }
    yb ^= 1;
  }
  return outs;
}
}  // namespace

int ttr_round_rand_tuple(ttr_handle_t h, int64_t n, const int64_t* nfactors, const int64_t* factor_ids,
                         const int64_t* rks, const ttr_rand_opts* opts, int64_t* out_ids) {
  return guarded(h, [&] {
    TTR_REQUIRE(out_ids && opts, ST_INVALID_ARG, "bad arguments");
    SketchPlan sp;
    auto its = tuple_items(h, n, nfactors, factor_ids, rks, opts, sp);
    auto outs = round_rand_tuple_batch(h, its, sp);
    for (int64_t i = 0; i < n; ++i) out_ids[i] = h->add(std::move(outs[size_t(i)]));
  });
}

namespace {
// Left-to-right recursive sketch of the entrywise product of the factors f, never formed
// (tt_recursive_sketch(T, A::NTuple, rks; reverse=false), src/sketches.jl:539-604), FP64, into
// host arrays.  Per core k the blocks j < nb of W_k (rows: the factors' left rank indices,
// first factor fastest) are contracted with the factors' cores one factor at a time
// (contract_sketch_core_kronecker_forwards!, src/sketches.jl:1332-1478): factor 0 for every
// (mode index z, block j) as one GEMM over all remaining columns, factor m >= 1 for every
// (remaining column, z, j) as one GEMM of the already-contracted indices against A_m[z].
// Intermediates are laid out [contracted, uncontracted, beta, z, j], so the last factor
// leaves T_j = [(a_1..a_M), (beta, z)] with beta fastest, the row order of the forward
// blocks S_j ((bv_k n_k) x bv_{k+1}), and W_{k+1}[:, block j] = T_j S_j is one GEMM per block.
// W_0 = 1 serves every block of core 0.  Cost n_k sum_m (rank products) per block and core
// instead of n_k P_k P_{k+1}.
void sketch_tuple_fwd(ttr_handle_s* h, const std::vector<DevTT*>& f, const int64_t* rks, const ttr_rand_opts* opts,
                      double* const* W_out, int64_t* sketch_rks) {
  Ctx& ctx = h->ctx;
  TTR_REQUIRE(rks != nullptr, ST_INVALID_ARG, "rks must not be NULL");
  TTR_REQUIRE(to_prec(opts->sketch_prec) == Prec::F64, ST_INVALID_ARG, "the forward sketch is FP64");
  const int d = f[0]->d;
  const size_t M = f.size();
  TTR_REQUIRE(rks[0] == 1, ST_INVALID_ARG, "rks[0] must be 1");
  for (auto* x : f) TTR_REQUIRE(x->rks[0] == 1 && x->rks[size_t(d)] == 1, ST_INVALID_ARG,
                                "factors need rks[0] = rks[d] = 1");
  auto P = [&](int k) {
    int64_t v = 1;
    for (auto* x : f) v *= x->rks[size_t(k)];
    return v;
  };
  for (int k = 0; k <= d; ++k) TTR_REQUIRE(P(k) < (int64_t(1) << 31), ST_INVALID_ARG, "product rank too large");
  FwdPlan fp;
  prepare_fwd(opts, d, f[0]->dims, fp);
  const std::vector<int64_t> r(rks, rks + d + 1);
  const std::vector<int64_t> p = block_counts_fwd(r, fp.bv, d, opts->p, opts->oversampling);
  fp.pmax = p;
  make_sketch_fwd(h, fp, d);
  // |T^(m+1)| = (prod_{q<=m} R_q) (prod_{q>m} L_q) bv_k n_k nb: the largest over cores and factors
  int64_t maxU = 1, maxT = 1;
  for (int k = 0; k < d; ++k) {
    const int64_t nb = k == 0 ? 1 : p[size_t(k) + 1], tail = fp.bv[size_t(k)] * fp.dims[size_t(k)] * nb;
    for (size_t m = 0; m < M; ++m) {
      int64_t v = tail;
      for (size_t q = 0; q <= m; ++q) v *= f[q]->rks[size_t(k) + 1];
      for (size_t q = m + 1; q < M; ++q) v *= f[q]->rks[size_t(k)];
      if (m + 1 == M) maxT = std::max(maxT, v);
      else maxU = std::max(maxU, v);
    }
  }
  std::vector<double*> W(size_t(d) + 1, nullptr);
  for (int k = 1; k <= d; ++k) W[size_t(k)] = ctx.ws.alloc_n<double>(size_t(P(k) * fp.bv[size_t(k)] * p[size_t(k)]));
  double* one = ctx.ws.alloc_n<double>(1);
  double* U[2] = {ctx.ws.alloc_n<double>(size_t(maxU)), ctx.ws.alloc_n<double>(size_t(maxU))};
  double* TM = ctx.ws.alloc_n<double>(size_t(maxT));
  fill2d_batched(ctx, {Fill2D{one, 1, 1, 1, 1.0}});
  GemmBatch gb;
  for (int k = 0; k < d; ++k) {
    const int64_t n = fp.dims[size_t(k)], bl = fp.bv[size_t(k)], br = fp.bv[size_t(k) + 1];
    const int64_t nb = k == 0 ? 1 : p[size_t(k) + 1];
    const double* V = k == 0 ? one : W[size_t(k)];
    const int64_t Pk = P(k), Pk1 = P(k + 1);
    int64_t pre = 1;  // prod of the right ranks of the factors already contracted
    for (size_t m = 0; m < M; ++m) {
      const DevTT* x = f[m];
      const int64_t L = x->rks[size_t(k)], R = x->rks[size_t(k) + 1];
      int64_t mid = 1;  // left ranks of the factors not yet contracted
      for (size_t q = m + 1; q < M; ++q) mid *= f[q]->rks[size_t(k)];
      const int64_t MB = mid * bl;
      const double* src = m == 0 ? V : U[(m - 1) & 1];
      double* dst = m + 1 == M ? TM : U[m & 1];
      for (int64_t j = 0; j < nb; ++j)
        for (int64_t z = 0; z < n; ++z) {
          const double* Az = x->cores[size_t(k)] + L * z;  // A_m[z](al, a) = core[al + L (z + n a)]
          const int64_t zj = z + n * j;
          if (m == 0) {  // [a_0, (mid, beta)] = A_0[z]^T V_j, V_j = L x MB (ld L)
            gb.add(int(R), int(MB), int(L), 1.0, Az, L * n, 1, src + Pk * bl * j, 1, L, 0.0, dst + R * MB * zj, 1, R);
            continue;
          }
          for (int64_t c = 0; c < MB; ++c)  // slab [pre, al_m] (ld pre) times A_m[z]
            gb.add(int(pre), int(R), int(L), 1.0, src + pre * L * (c + MB * zj), 1, pre, Az, 1, L * n, 0.0,
                   dst + pre * R * (c + MB * zj), 1, pre);
        }
      gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "tuple_fwd_cores");
      pre *= R;
    }
    // W_{k+1}[:, block j] = T_j S_j over (beta, z), T_j = TM block j (P_{k+1} x bv_k n)
    for (int64_t j = 0; j < p[size_t(k) + 1]; ++j)
      gb.add(int(Pk1), int(br), int(bl * n), 1.0, TM + Pk1 * bl * n * (k == 0 ? 0 : j), 1, Pk1,
             fp.S[size_t(k)] + bl * n * br * j, 1, bl * n, 0.0, W[size_t(k) + 1] + Pk1 * br * j, 1, Pk1);
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "tuple_fwd_w");
  }
  ctx.sync();
  for (int k = 0; k <= d; ++k) {
    const int64_t sk = fp.bv[size_t(k)] * p[size_t(k)];
    if (sketch_rks) sketch_rks[k] = sk;
    if (!W_out || !W_out[k]) continue;
    const int64_t cnt = P(k) * sk;
    const double scale = 1.0 / std::sqrt(double(p[size_t(k)]));
    if (k == 0) {
      for (int64_t e = 0; e < cnt; ++e) W_out[k][e] = scale;
      continue;
    }
    TTR_HIP(hipMemcpy(W_out[k], W[size_t(k)], size_t(cnt) * 8, hipMemcpyDeviceToHost));
    for (int64_t e = 0; e < cnt; ++e) W_out[k][e] *= scale;
  }
}
}  // namespace

int ttr_sketch_tuple(ttr_handle_t h, int64_t m, const int64_t* factor_ids, const int64_t* rks,
                     const ttr_rand_opts* opts, double* const* W_out, int64_t* sketch_rks) {
  return guarded(h, [&] {
    TTR_REQUIRE(opts, ST_INVALID_ARG, "options must not be NULL");
    if (opts->forward) {
      TTR_REQUIRE(m >= 1 && m <= KRON_MAX && factor_ids, ST_INVALID_ARG, "a product needs 1..8 factors");
      std::vector<DevTT*> f;
      for (int64_t q = 0; q < m; ++q) {
        f.push_back(h->get(factor_ids[q]));
        check_dims_equal(f[0], f.back());
      }
      sketch_tuple_fwd(h, f, rks, opts, W_out, sketch_rks);
      return;
    }
    SketchPlan sp;
    const int64_t nf = m;
    auto its = tuple_items(h, 1, &nf, factor_ids, rks, opts, sp);
    const int d = sp.d;
    make_sketch(h, sp, 0);
    tuple_sketch_phase(h, sp, its, 0);
    h->ctx.sync();
    const TupleItem& it = its[0];
    std::vector<int32_t> E(static_cast<size_t>(d + 1));
    TTR_HIP(hipMemcpy(E.data(), it.E, size_t(d + 1) * 4, hipMemcpyDeviceToHost));
    const size_t eb = prec_bytes(sp.wprec);
    for (int k = 0; k <= d; ++k) {
      if (sketch_rks) sketch_rks[k] = it.s[size_t(k)];
      if (!W_out || !W_out[k]) continue;
      const int64_t cnt = it.P(k) * it.s[size_t(k)];
      const double scale = std::ldexp(1.0, k == d ? 0 : E[size_t(k)]) / std::sqrt(double(it.p[size_t(k)]));
      if (k == d) {
        for (int64_t e = 0; e < cnt; ++e) W_out[k][e] = scale;
        continue;
      }
      std::vector<char> raw(size_t(cnt) * eb);
      TTR_HIP(hipMemcpy(raw.data(), it.W[size_t(k)], raw.size(), hipMemcpyDeviceToHost));
      for (int64_t e = 0; e < cnt; ++e) {
        double v;
        if (sp.wprec == Prec::F64) std::memcpy(&v, raw.data() + 8 * e, 8);
        else if (sp.wprec == Prec::F32) { float f; std::memcpy(&f, raw.data() + 4 * e, 4); v = f; }
        else { uint16_t b; std::memcpy(&b, raw.data() + 2 * e, 2); v = bf16_bits_to_float(b); }
        W_out[k][e] = v * scale;
      }
    }
  });
}

int ttr_round_rand(ttr_handle_t h, int64_t n, const ttr_rand_item* items, const ttr_rand_opts* opts,
                   int64_t* out_ids) {
  return guarded(h, [&] {
    TTR_REQUIRE(out_ids, ST_INVALID_ARG, "bad arguments");
    std::vector<RandItem> its = parse_items(h, n, items);
    auto outs = round_rand_graphed(h, its, opts);
    for (int64_t i = 0; i < n; ++i) out_ids[i] = h->add(std::move(outs[i]));
  });
}

namespace {
// reverse recursive sketch of x (or of H x without forming it) into host arrays
void sketch_one(ttr_handle_s* h, DevTT* x, DevTTO* H, const int64_t* rks, const ttr_rand_opts* opts,
                double* const* W_out, int64_t* sketch_rks) {
  TTR_REQUIRE(rks != nullptr, ST_INVALID_ARG, "rks must not be NULL");
  const int d = x->d;
  TTR_REQUIRE(x->rks[d] == 1, ST_INVALID_ARG, "rks[d] must be 1");
  if (H) TTR_REQUIRE(H->dims == x->dims, ST_DIM_MISMATCH, "operator and TT dims differ");
  std::vector<RandItem> its(1);
  RandTerm t{};
  t.x = x;
  t.H = H;
  t.alpha = 1.0;
  its[0].terms.push_back(t);
  its[0].rks.assign(rks, rks + d + 1);
  SketchPlan sp;
  prepare_plan(h, opts, d, x->dims, its, sp);
  make_sketch(h, sp, 0);
  sketch_phase(h, sp, its, 0);
  h->ctx.sync();
  const RandTerm& tt = its[0].terms[0];
  std::vector<int32_t> E(static_cast<size_t>(d + 1));
  TTR_HIP(hipMemcpy(E.data(), tt.E, size_t(d + 1) * 4, hipMemcpyDeviceToHost));
  const size_t eb = prec_bytes(sp.wprec);
  for (int k = 0; k <= d; ++k) {
    if (sketch_rks) sketch_rks[k] = its[0].s[k];
    if (!W_out || !W_out[k]) continue;
    const int64_t cnt = tt.rc(k) * its[0].s[k];
    const double scale = std::ldexp(1.0, k == d ? 0 : E[k]) / std::sqrt(double(its[0].p[k]));
    if (k == d) {
      for (int64_t e = 0; e < cnt; ++e) W_out[k][e] = 1.0 / std::sqrt(double(its[0].p[k]));
      continue;
    }
    std::vector<char> raw(size_t(cnt) * eb);
    TTR_HIP(hipMemcpy(raw.data(), tt.W[k], raw.size(), hipMemcpyDeviceToHost));
    for (int64_t e = 0; e < cnt; ++e) {
      double v;
      if (sp.wprec == Prec::F64) std::memcpy(&v, raw.data() + 8 * e, 8);
      else if (sp.wprec == Prec::F32) { float f; std::memcpy(&f, raw.data() + 4 * e, 4); v = f; }
      else { uint16_t b; std::memcpy(&b, raw.data() + 2 * e, 2); v = bf16_bits_to_float(b); }
      W_out[k][e] = v * scale;
    }
  }
}
}  // namespace

namespace {
// forward recursive sketch of one TT into host arrays (W_0 = 1)
void sketch_one_fwd(ttr_handle_s* h, DevTT* x, DevTTO* H, const int64_t* rks, const ttr_rand_opts* opts,
                    double* const* W_out, int64_t* sketch_rks) {
  TTR_REQUIRE(rks != nullptr, ST_INVALID_ARG, "rks must not be NULL");
  TTR_REQUIRE(to_prec(opts->sketch_prec) == Prec::F64, ST_INVALID_ARG, "the forward sketch is FP64");
  const int d = x->d;
  TTR_REQUIRE(rks[0] == 1 && x->rks[0] == 1, ST_INVALID_ARG, "rks[0] must be 1");
  if (H) {
    TTR_REQUIRE(H->dims == x->dims, ST_DIM_MISMATCH, "operator and TT dims differ");
    TTR_REQUIRE(H->rks[0] == 1 && H->rks[d] == 1, ST_INVALID_ARG, "operator boundary ranks must be 1");
  }
  FwdPlan fp;
  prepare_fwd(opts, d, x->dims, fp);
  std::vector<int64_t> r(rks, rks + d + 1);
  FwdTerm t;
  t.x = x;
  t.H = H;
  // the operator form takes no oversampling (src/sketches.jl:399-402)
  t.p = block_counts_fwd(r, fp.bv, d, opts->p, H ? 1.0 : opts->oversampling);
  fp.pmax = t.p;
  std::vector<FwdTerm> ts{t};
  make_sketch_fwd(h, fp, d);
  sketch_phase_fwd(h, fp, ts, d);
  h->ctx.sync();
  for (int k = 0; k <= d; ++k) {
    const int64_t s = fp.bv[k] * ts[0].p[k];
    if (sketch_rks) sketch_rks[k] = s;
    if (!W_out || !W_out[k]) continue;
    const int64_t cnt = ts[0].rc(k) * s;
    const double scale = 1.0 / std::sqrt(double(ts[0].p[k]));
    if (k == 0) {
      for (int64_t e = 0; e < cnt; ++e) W_out[k][e] = scale;
      continue;
    }
    TTR_HIP(hipMemcpy(W_out[k], ts[0].W[k], size_t(cnt) * 8, hipMemcpyDeviceToHost));
    for (int64_t e = 0; e < cnt; ++e) W_out[k][e] *= scale;
  }
}
}  // namespace

int ttr_sketch(ttr_handle_t h, int64_t tt_id, const int64_t* rks, const ttr_rand_opts* opts, double* const* W_out,
               int64_t* sketch_rks) {
  return guarded(h, [&] {
    TTR_REQUIRE(opts != nullptr, ST_INVALID_ARG, "options must not be NULL");
    if (opts->forward) sketch_one_fwd(h, h->get(tt_id), nullptr, rks, opts, W_out, sketch_rks);
    else sketch_one(h, h->get(tt_id), nullptr, rks, opts, W_out, sketch_rks);
  });
}

int ttr_stta(ttr_handle_t h, int64_t n, const ttr_rand_item* items, const ttr_rand_opts* opts, uint64_t seed_right,
             int64_t* out_ids) {
  return guarded(h, [&] {
    TTR_REQUIRE(out_ids, ST_INVALID_ARG, "bad arguments");
    std::vector<SttaItem> its = parse_stta_items(h, n, items);
    stta_sketch_batch(h, its, opts, seed_right);
    auto outs = stta_solve_batch(h, its);
    for (int64_t i = 0; i < n; ++i) out_ids[i] = h->add(std::move(outs[size_t(i)]));
  });
}

int ttr_stta_sketch(ttr_handle_t h, int64_t tt_id, const int64_t* rks, const ttr_rand_opts* opts,
                    uint64_t seed_right, double* const* Omega_out, double* const* Psi_out, int64_t* sl,
                    int64_t* sr) {
  return guarded(h, [&] {
    TTR_REQUIRE(rks != nullptr, ST_INVALID_ARG, "rks must not be NULL");
    DevTT* x = h->get(tt_id);
    const int d = x->d;
    const double one = 1.0;
    ttr_rand_item it{1, &one, &tt_id, -1, rks};
    std::vector<SttaItem> its = parse_stta_items(h, 1, &it);
    stta_sketch_batch(h, its, opts, seed_right);
    h->ctx.sync();
    const SttaItem& S = its[0];
    for (int k = 0; k <= d; ++k) {
      if (sl) sl[k] = S.sl[k];
      if (sr) sr[k] = S.sr[k];
    }
    for (int k = 0; k + 1 < d; ++k)
      if (Omega_out && Omega_out[k])
        TTR_HIP(hipMemcpy(Omega_out[k], S.Om[k], size_t(S.sl[k + 1] * S.sr[k + 1]) * 8, hipMemcpyDeviceToHost));
    for (int k = 0; k < d; ++k) {
      if (!Psi_out || !Psi_out[k]) continue;
      const int64_t a = S.sl[k], nk = x->dims[k], b = S.sr[k + 1];
      std::vector<double> raw(size_t(a * nk * b));
      TTR_HIP(hipMemcpy(raw.data(), S.Psi[k], raw.size() * 8, hipMemcpyDeviceToHost));
      // device (a, n, b) -> reference (n, a, b)
      for (int64_t bb = 0; bb < b; ++bb)
        for (int64_t z = 0; z < nk; ++z)
          for (int64_t aa = 0; aa < a; ++aa) Psi_out[k][z + nk * (aa + a * bb)] = raw[size_t(aa + a * (z + nk * bb))];
    }
  });
}

// stta_sketch(x, L, R) with explicit sketch TTs (src/sketches.jl:1112-1124): the
// left environments of (x, L) and the right ones of (x, R) (partial_contraction,
// batched GEMMs), then Omega_k = WL_{k+1}^T WR_{k+1} and
// Psi_k = WL_k^T X_k WR_{k+1}, all on the device.
int ttr_stta_sketch_tt(ttr_handle_t h, int64_t x_id, int64_t l_id, int64_t r_id, double* const* Omega_out,
                       double* const* Psi_out) {
  return guarded(h, [&] {
    DevTT* x = h->get(x_id);
    DevTT* L = h->get(l_id);
    DevTT* R = h->get(r_id);
    TTR_REQUIRE(x->dims == L->dims && x->dims == R->dims, ST_DIM_MISMATCH, "TT dimensions are not compatible");
    const int d = x->d;
    Ctx& ctx = h->ctx;
    std::vector<std::vector<double*>> WL(1, std::vector<double*>(size_t(d + 1))), WR = WL;
    for (int k = 0; k <= d; ++k) {
      WL[0][size_t(k)] = k == 0 ? h->one : ctx.ws.alloc_n<double>(size_t(x->rks[k] * L->rks[k]));
      WR[0][size_t(k)] = k == d ? h->one : ctx.ws.alloc_n<double>(size_t(x->rks[k] * R->rks[k]));
    }
    pair_environments(h, {x}, {L}, false, WL);
    pair_environments(h, {x}, {R}, true, WR);
    std::vector<double*> Om(static_cast<size_t>(std::max(d - 1, 1))), Ps(static_cast<size_t>(d)),
        T(static_cast<size_t>(d));
    GemmBatch gb;
    for (int k = 0; k + 1 < d; ++k) {  // Omega_k (rL_{k+1} x rR_{k+1})
      const int64_t z = x->rks[k + 1], a = L->rks[k + 1], b = R->rks[k + 1];
      Om[size_t(k)] = ctx.ws.alloc_n<double>(size_t(a * b));
      gb.add(int(a), int(b), int(z), 1.0, WL[0][size_t(k + 1)], z, 1, WR[0][size_t(k + 1)], 1, z, 0.0, Om[size_t(k)], 1,
             a);
    }
    for (int k = 0; k < d; ++k) {  // T = WL_k^T X_k: (rL_k) x (n r_{k+1})
      const int64_t y = x->rks[k], n = x->dims[k], z1 = x->rks[k + 1], a = L->rks[k];
      T[size_t(k)] = ctx.ws.alloc_n<double>(size_t(a * n * z1));
      gb.add(int(a), int(n * z1), int(y), 1.0, WL[0][size_t(k)], y, 1, x->cores[k], 1, y, 0.0, T[size_t(k)], 1, a);
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "stta_tt_a");
    for (int k = 0; k < d; ++k) {  // Psi_k = T (rL_k n) x r_{k+1} times WR_{k+1}: device layout (rL_k, n, rR_{k+1})
      const int64_t n = x->dims[k], z1 = x->rks[k + 1], a = L->rks[k], b = R->rks[k + 1];
      Ps[size_t(k)] = ctx.ws.alloc_n<double>(size_t(a * n * b));
      gb.add(int(a * n), int(b), int(z1), 1.0, T[size_t(k)], 1, a * n, WR[0][size_t(k + 1)], 1, z1, 0.0,
             Ps[size_t(k)], 1, a * n);
    }
    gemm_run(ctx, gb, Prec::F64, Prec::F64, Prec::F64, Prec::F64, "stta_tt_b");
    ctx.sync();
    for (int k = 0; k + 1 < d; ++k)
      if (Omega_out && Omega_out[k])
        TTR_HIP(hipMemcpy(Omega_out[k], Om[size_t(k)], size_t(L->rks[k + 1] * R->rks[k + 1]) * 8,
                          hipMemcpyDeviceToHost));
    for (int k = 0; k < d; ++k) {
      if (!Psi_out || !Psi_out[k]) continue;
      const int64_t a = L->rks[k], nk = x->dims[k], b = R->rks[k + 1];
      std::vector<double> raw(size_t(a * nk * b));
      TTR_HIP(hipMemcpy(raw.data(), Ps[size_t(k)], raw.size() * 8, hipMemcpyDeviceToHost));
      for (int64_t bb = 0; bb < b; ++bb)  // device (a, n, b) -> reference (n, a, b)
        for (int64_t zz = 0; zz < nk; ++zz)
          for (int64_t aa = 0; aa < a; ++aa) Psi_out[k][zz + nk * (aa + a * bb)] = raw[size_t(aa + a * (zz + nk * bb))];
    }
  });
}

int ttr_sketch_op(ttr_handle_t h, int64_t op_id, int64_t tt_id, const int64_t* rks, const ttr_rand_opts* opts,
                  double* const* W_out, int64_t* sketch_rks) {
  return guarded(h, [&] {
    TTR_REQUIRE(opts != nullptr, ST_INVALID_ARG, "options must not be NULL");
    if (opts->forward) sketch_one_fwd(h, h->get(tt_id), h->get_op(op_id), rks, opts, W_out, sketch_rks);
    else sketch_one(h, h->get(tt_id), h->get_op(op_id), rks, opts, W_out, sketch_rks);
  });
}

// tt_sketch(T, [H,] A, s; reverse=true, ...) (src/sketches.jl:645-723, operator
// form :729-813) for a batch: the right-to-left recursion of the recursive sketch
// with the same block count p = cld(s, bv_1) at every core (the draws of core k
// are those of the recursive sketch with p blocks), all items in the same
// launches; only the boundary sketch vec(W_1) / sqrt(p) is returned.
int ttr_tt_sketch(ttr_handle_t h, int64_t n, const int64_t* tt_ids, const int64_t* op_ids, int64_t s,
                  const ttr_rand_opts* opts, double* out, int64_t out_len, int64_t* s_actual) {
  return guarded(h, [&] {
    TTR_REQUIRE(n >= 1 && tt_ids && opts && s_actual && s >= 1, ST_INVALID_ARG, "bad arguments");
    DevTT* x0 = h->get(tt_ids[0]);
    const int d = x0->d;
    const std::vector<int64_t> bv = block_ranks(x0->dims, opts->block_rks > 0 ? opts->block_rks : d);
    const int64_t p = (s + bv[0] - 1) / bv[0];
    const int64_t sa = bv[0] * p;
    *s_actual = sa;
    if (!out) return;  // size query
    TTR_REQUIRE(out_len >= n * sa, ST_INVALID_ARG, "out holds fewer than n * s_actual values");
    std::vector<RandItem> its(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
      RandTerm t{};
      t.x = h->get(tt_ids[i]);
      t.alpha = 1.0;
      t.H = (op_ids && op_ids[i] >= 0) ? h->get_op(op_ids[i]) : nullptr;
      check_dims_equal(x0, t.x);
      TTR_REQUIRE(t.x->rks[0] == 1 && t.x->rks[d] == 1, ST_INVALID_ARG, "boundary ranks must be 1");
      if (t.H) {
        TTR_REQUIRE(t.H->dims == x0->dims, ST_DIM_MISMATCH, "operator and TT dims differ");
        TTR_REQUIRE(t.H->rks[0] == 1 && t.H->rks[d] == 1, ST_INVALID_ARG, "operator boundary ranks must be 1");
      }
      its[size_t(i)].terms.push_back(t);
      its[size_t(i)].rks.assign(size_t(d) + 1, 1);
    }
    std::vector<int64_t> pv(size_t(d) + 1, p);
    pv[size_t(d)] = 1;
    ttr_rand_opts o2 = *opts;
    o2.p = pv.data();
    o2.oversampling = 1.0;
    SketchPlan sp;
    prepare_plan(h, &o2, d, x0->dims, its, sp);
    make_sketch(h, sp, 0);
    sketch_phase(h, sp, its, 0);
    Ctx& ctx = h->ctx;
    const size_t eb = prec_bytes(sp.wprec);
    // W_1 (1 x s) and its power-of-two exponent of every item, one sync
    char* hw = static_cast<char*>(ctx.pin.alloc(size_t(n) * size_t(sa) * eb));
    int32_t* he = static_cast<int32_t*>(ctx.pin.alloc(sizeof(int32_t) * size_t(n)));
    for (int64_t i = 0; i < n; ++i) {
      const RandTerm& tt = its[size_t(i)].terms[0];
      TTR_HIP(hipMemcpyAsync(hw + size_t(i) * size_t(sa) * eb, tt.W[0], size_t(sa) * eb, hipMemcpyDeviceToHost,
                             ctx.stream));
      TTR_HIP(hipMemcpyAsync(he + i, tt.E, sizeof(int32_t), hipMemcpyDeviceToHost, ctx.stream));
    }
    ctx.sync();
    for (int64_t i = 0; i < n; ++i) {
      const double scale = std::ldexp(1.0, he[i]) / std::sqrt(double(p));
      const char* raw = hw + size_t(i) * size_t(sa) * eb;
      for (int64_t e = 0; e < sa; ++e) {
        double v;
        if (sp.wprec == Prec::F64) std::memcpy(&v, raw + 8 * e, 8);
        else if (sp.wprec == Prec::F32) { float f; std::memcpy(&f, raw + 4 * e, 4); v = f; }
        else { uint16_t b; std::memcpy(&b, raw + 2 * e, 2); v = bf16_bits_to_float(b); }
        out[i * sa + e] = v * scale;
      }
    }
  });
}

int ttr_sketch_blocks(ttr_handle_t h, int64_t d, const int64_t* dims, const int64_t* rks, const ttr_rand_opts* opts,
                      int64_t k, double* out, int64_t* shape4) {
  return guarded(h, [&] {
    TTR_REQUIRE(d >= 1 && dims && rks && out && k >= 0 && k < d, ST_INVALID_ARG, "bad arguments");
    std::vector<int64_t> vd(dims, dims + d);
    if (opts && opts->forward) {
      TTR_REQUIRE(to_prec(opts->sketch_prec) == Prec::F64, ST_INVALID_ARG, "the forward sketch is FP64");
      FwdPlan fp;
      prepare_fwd(opts, int(d), vd, fp);
      std::vector<int64_t> r(rks, rks + d + 1);
      fp.pmax = block_counts_fwd(r, fp.bv, int(d), opts->p, opts->oversampling);
      make_sketch_fwd(h, fp, int(k) + 1);
      h->ctx.sync();
      const int64_t left = fp.bv[k], right = fp.bv[k + 1], p = fp.pmax[k + 1], nk = vd[k];
      TTR_HIP(hipMemcpy(out, fp.S[k], size_t(left * nk * right * p) * 8, hipMemcpyDeviceToHost));
      if (shape4) {
        shape4[0] = left;
        shape4[1] = nk;
        shape4[2] = right;
        shape4[3] = p;
      }
      return;
    }
    std::vector<RandItem> its(1);
    its[0].rks.assign(rks, rks + d + 1);
    SketchPlan sp;
    prepare_plan(h, opts, int(d), vd, its, sp);
    sp.pmax = its[0].p;
    make_sketch(h, sp, int(k));
    h->ctx.sync();
    const int64_t n = vd[k], bvr = sp.bv[k + 1], bvl = sp.bv[k], p = sp.pmax[k];
    const int64_t cnt = n * bvr * bvl * p;
    const size_t eb = prec_bytes(sp.prec);
    std::vector<char> raw(size_t(cnt) * eb);
    TTR_HIP(hipMemcpy(raw.data(), sp.S[k], raw.size(), hipMemcpyDeviceToHost));
    for (int64_t e = 0; e < cnt; ++e) {
      const int64_t beta = e % bvr, rest = e / bvr, z = rest % n, col = rest / n;
      const int64_t src = z + n * (beta + bvr * col);
      double v;
      if (sp.prec == Prec::F64) std::memcpy(&v, raw.data() + 8 * src, 8);
      else if (sp.prec == Prec::F32) { float f; std::memcpy(&f, raw.data() + 4 * src, 4); v = f; }
      else { uint16_t b; std::memcpy(&b, raw.data() + 2 * src, 2); v = bf16_bits_to_float(b); }
      out[e] = v;
    }
    if (shape4) {
      shape4[0] = bvr;
      shape4[1] = n;
      shape4[2] = bvl;
      shape4[3] = p;
    }
  });
}

int ttr_graph_enable(ttr_handle_t h, int on) {
  return guarded(h, [&] {
    h->graphs = on != 0;
    if (!h->graphs) h->graph_clear();
  });
}

int ttr_graph_stats(ttr_handle_t h, int64_t* captures, int64_t* replays, int64_t* fallbacks, int64_t* cached) {
  return guarded(h, [&] {
    if (captures) *captures = h->g_captures;
    if (replays) *replays = h->g_replays;
    if (fallbacks) *fallbacks = h->g_fallbacks;
    if (cached) {
      int64_t c = 0;
      for (auto& kv : h->gcache) c += kv.second.state == 1;
      *cached = c;
    }
  });
}

int ttr_trim(ttr_handle_t h) {
  return guarded(h, [&] {
    TTR_HIP(hipStreamSynchronize(h->ctx.stream));
    h->ctx.ws.release();
    h->ctx.cache.release_cached();
  });
}

int ttr_profile_enable(ttr_handle_t h, int on) {
  return guarded(h, [&] {
    TTR_REQUIRE(on >= 0 && on <= 2, ST_INVALID_ARG, "profiling level must be 0, 1 or 2");
    h->ctx.prof.on = on != 0;
    h->ctx.prof.level = on == 2 ? 2 : 1;
  });
}

int ttr_profile_reset(ttr_handle_t h) {
  return guarded(h, [&] { h->ctx.prof.reset(); });
}

int ttr_profile_count(ttr_handle_t h, int64_t* count) {
  return guarded(h, [&] {
    TTR_REQUIRE(count, ST_INVALID_ARG, "null count");
    *count = int64_t(h->ctx.prof.stats.size());
  });
}

int ttr_profile_get(ttr_handle_t h, int64_t i, char* name, int64_t name_len, double* ms, int64_t* launches,
                    double* flops, double* bytes) {
  return guarded(h, [&] {
    auto& st = h->ctx.prof.stats;
    TTR_REQUIRE(i >= 0 && i < int64_t(st.size()), ST_INVALID_ARG, "profile index out of range");
    auto it = st.begin();
    std::advance(it, i);
    if (name && name_len > 0) {
      std::strncpy(name, it->first.c_str(), size_t(name_len - 1));
      name[name_len - 1] = 0;
    }
    if (ms) *ms = it->second.ms;
    if (launches) *launches = it->second.launches;
    if (flops) *flops = it->second.flops;
    if (bytes) *bytes = it->second.bytes;
  });
}

int ttr_sketch_draws(ttr_handle_t h, int64_t k, int64_t count, uint64_t seed, double* out) {
  return guarded(h, [&] {
    TTR_REQUIRE(count >= 0 && out && k >= 0, ST_INVALID_ARG, "bad arguments");
    double* d = h->ctx.ws.alloc_n<double>(size_t(std::max<int64_t>(count, 1)));
    sketch_draws(h->ctx, d, count, int(k), seed);
    h->ctx.sync();
    TTR_HIP(hipMemcpy(out, d, size_t(count) * 8, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
