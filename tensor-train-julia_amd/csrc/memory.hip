// Device and pinned-host memory pools of libttround.
#include "ttr_internal.h"

namespace ttr {

namespace {
constexpr size_t kAlign = 256;
constexpr size_t kMinBlock = size_t(64) << 20;
inline size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }
}  // namespace

DevArena::~DevArena() { release(); }

void* DevArena::alloc(size_t bytes) {
  bytes = align_up(bytes ? bytes : kAlign);
  if (fixed_) {
    TTR_REQUIRE(fixed_used_ + bytes <= fixed_cap_, ST_INVALID_ARG,
                "the call needs more than the " + std::to_string(fixed_cap_ >> 20) +
                    " MiB of device scratch of the non-allocating entry points");
    void* p = fixed_ + fixed_used_;
    fixed_used_ += bytes;
    fixed_peak_ = std::max(fixed_peak_, fixed_used_);
    return p;
  }
  // the current block, else a later one (blocks kept after a rewind), else a new one
  for (; cur_ < blocks_.size(); ++cur_) {
    Block& b = blocks_[cur_];
    if (b.used + bytes <= b.cap) {
      void* p = b.p + b.used;
      b.used += bytes;
      peak_ = std::max(peak_, held());
      return p;
    }
  }
  size_t cap = bytes > kMinBlock ? bytes : kMinBlock;
  if (!blocks_.empty() && blocks_.back().cap > cap) cap = blocks_.back().cap;
  char* p = nullptr;
  TTR_HIP(hipMalloc(&p, cap));
  blocks_.push_back({p, cap, bytes});
  cur_ = blocks_.size() - 1;
  peak_ = std::max(peak_, held());
  return p;
}

DevArena::Mark DevArena::mark() const {
  if (fixed_) return {0, fixed_used_};
  if (cur_ >= blocks_.size()) return {blocks_.size(), 0};
  return {cur_, blocks_[cur_].used};
}

void DevArena::rewind(const Mark& m) {
  if (fixed_) {
    fixed_used_ = std::min(fixed_used_, m.off);
    return;
  }
  for (size_t i = m.blk + 1; i < blocks_.size(); ++i) blocks_[i].used = 0;
  if (m.blk < blocks_.size()) blocks_[m.blk].used = m.off;
  cur_ = std::min(m.blk, blocks_.empty() ? size_t(0) : blocks_.size() - 1);
}

size_t DevArena::held() const {
  size_t u = 0;
  for (auto& b : blocks_) u += b.used;
  return u;
}

void DevArena::reset() {
  cur_ = 0;
  peak_ = 0;
  if (fixed_) {
    fixed_used_ = 0;
    fixed_peak_ = 0;
    return;
  }
  if (blocks_.size() > 1) {
    size_t total = 0;
    for (auto& b : blocks_) total += b.cap;
    release();
    char* p = nullptr;
    TTR_HIP(hipMalloc(&p, total));
    blocks_.push_back({p, total, 0});
  } else if (!blocks_.empty()) {
    blocks_[0].used = 0;
  }
}

size_t DevArena::used() const { return fixed_ ? fixed_peak_ : peak_; }

void DevArena::release() {
  for (auto& b : blocks_) (void)hipFree(b.p);
  blocks_.clear();
  cur_ = 0;
}

PinnedArena::~PinnedArena() { release(); }

void* PinnedArena::alloc(size_t bytes) {
  bytes = align_up(bytes ? bytes : kAlign);
  if (fixed_) {
    TTR_REQUIRE(fixed_used_ + bytes <= fixed_cap_, ST_INVALID_ARG,
                "the call needs more than the " + std::to_string(fixed_cap_ >> 20) +
                    " MiB of host staging of the non-allocating entry points");
    void* p = fixed_ + fixed_used_;
    fixed_used_ += bytes;
    return p;
  }
  if (!blocks_.empty()) {
    Block& b = blocks_.back();
    if (b.used + bytes <= b.cap) {
      void* p = b.p + b.used;
      b.used += bytes;
      return p;
    }
  }
  size_t cap = bytes > (size_t(8) << 20) ? bytes : (size_t(8) << 20);
  if (!blocks_.empty() && blocks_.back().cap > cap) cap = blocks_.back().cap;
  char* p = nullptr;
  TTR_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), cap, hipHostMallocDefault));
  blocks_.push_back({p, cap, bytes});
  return p;
}

void PinnedArena::reset() {
  if (fixed_) {
    fixed_used_ = 0;
    return;
  }
  if (blocks_.size() > 1) {
    size_t total = 0;
    for (auto& b : blocks_) total += b.cap;
    release();
    char* p = nullptr;
    TTR_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), total, hipHostMallocDefault));
    blocks_.push_back({p, total, 0});
  } else if (!blocks_.empty()) {
    blocks_[0].used = 0;
  }
}

size_t PinnedArena::used() const {
  if (fixed_) return fixed_used_;
  size_t u = 0;
  for (auto& b : blocks_) u += b.used;
  return u;
}

void PinnedArena::release() {
  for (auto& b : blocks_) (void)hipHostFree(b.p);
  blocks_.clear();
}

CachingAlloc::~CachingAlloc() {
  release_cached();
  for (auto& kv : live_) (void)hipFree(kv.first);
  live_.clear();
}

size_t CachingAlloc::round_size(size_t b) {
  if (b < (size_t(1) << 20)) return align_up(b < 512 ? 512 : b);
  // round to 1/8 of the enclosing power of two
  size_t p = 1;
  while (p < b) p <<= 1;
  size_t step = p >> 3;
  return (b + step - 1) / step * step;
}

void* CachingAlloc::alloc(size_t bytes) {
  size_t sz = round_size(bytes);
  auto it = free_.find(sz);
  void* p = nullptr;
  if (it != free_.end()) {
    p = it->second;
    free_.erase(it);
  } else {
    hipError_t e = hipMalloc(&p, sz);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      release_cached();
      e = hipMalloc(&p, sz);
    }
    TTR_HIP(e);
  }
  live_[p] = sz;
  return p;
}

void CachingAlloc::free(void* p) {
  if (!p) return;
  auto it = live_.find(p);
  if (it == live_.end()) return;
  free_.emplace(it->second, p);
  live_.erase(it);
}

void CachingAlloc::release_cached() {
  for (auto& kv : free_) (void)hipFree(kv.second);
  free_.clear();
}

Profiler::~Profiler() {
  for (auto& kv : pending_) {
    pool_.push_back(kv.second.a);
    pool_.push_back(kv.second.b);
  }
  for (auto e : pool_) (void)hipEventDestroy(e);
}

hipEvent_t Profiler::get_event() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  TTR_HIP(hipEventCreate(&e));
  return e;
}

int Profiler::start(hipStream_t s, const char* name, double flops, double bytes, const char* name2) {
  // bound the number of outstanding event records: drain the closed ones first
  if (stopped_ >= 1024) {
    TTR_HIP(hipStreamSynchronize(s));
    collect();
  }
  Pending p{get_event(), get_event(), name, name2 ? name2 : "", flops, bytes, false};
  TTR_HIP(hipEventRecord(p.a, s));
  pending_.emplace(next_, p);
  return next_++;
}

void Profiler::stop(hipStream_t s, int token) {
  auto it = pending_.find(token);
  if (it == pending_.end() || it->second.stopped) return;
  (void)hipEventRecord(it->second.b, s);
  it->second.stopped = true;
  ++stopped_;
}

void Profiler::collect() {
  for (auto it = pending_.begin(); it != pending_.end();) {
    Pending& p = it->second;
    if (!p.stopped) {
      ++it;
      continue;
    }
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      for (const std::string* nm : {&p.name, &p.name2}) {
        if (nm->empty()) continue;
        ProfStat& st = stats[*nm];
        st.ms += ms;
        st.launches += 1;
        st.flops += p.flops;
        st.bytes += p.bytes;
      }
    } else {
      (void)hipGetLastError();
    }
    pool_.push_back(p.a);
    pool_.push_back(p.b);
    it = pending_.erase(it);
  }
  stopped_ = 0;
}

}  // namespace ttr

namespace ttr {

double* Ctx::splitk_buffer(size_t elems) {
  splitk_peak = std::max(splitk_peak, elems);
  if (ws.fixed()) {
    // one region reused by every launch of the call (they run in stream order)
    if (elems <= fsplit_elems) return fsplit;
    if (elems * sizeof(double) + 256 > ws.fixed_free()) return nullptr;
    fsplit = ws.alloc_n<double>(elems);
    fsplit_elems = elems;
    return fsplit;
  }
  if (elems > splitk_elems) {
    // the previous buffer may still be read by queued launches
    sync();
    if (splitk) TTR_HIP(hipFree(splitk));
    splitk = nullptr;
    splitk_elems = 0;
    const size_t n = std::max(elems, size_t(1) << 20);
    TTR_HIP(hipMalloc(reinterpret_cast<void**>(&splitk), n * sizeof(double)));
    splitk_elems = n;
  }
  return splitk;
}

}  // namespace ttr
