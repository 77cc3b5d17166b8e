// Serving loop: continuous batching of solve requests over resident problem types.
//
// The reference solves one problem instance per synchronous launch
// (GPU_HC_Solver::Solve_by_GPU_HC, GPU-HC/magmaHC/GPU_HC_Solver.cpp:207-384;
// its RANSAC loop re-runs that per target, cmd/magmaHC-main.cpp:93-112).
// Here requests (one parameter instance of a registered type each) are
// queued per type; a scheduler thread forms micro-batches, admits them while
// the HBM budget allows, and launches each on one of up to four streams, so
// work of requests that arrive while others are on the GPU overlaps on the
// device.  Tables and start roots stay resident in the problem objects; only
// the target parameters travel per request (pinned staging), and P2C runs on
// the device.  Every path is tracked with the same arithmetic whatever batch
// it lands in (one single-type launch per micro-batch, same kernel instance
// and launch shape), so a request's result does not depend on the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "hcx_internal.hpp"
#include "hcx_problem.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int type = 0;
  std::vector<hcx_complex> params;
  Clock::time_point t_submit, t_launch, t_done;
  bool done = false;
  std::vector<hcx_complex> roots;
  std::vector<uint32_t> flags;
  int refined = 0, real = 0;
};

// First-fit allocator over one device allocation (the HBM budget).
struct Arena {
  char* base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> length
  void init(char* b, size_t n) {
    base = b;
    size = n;
    free_.clear();
    free_[0] = n;
  }
  bool alloc(size_t n, size_t& off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        off = it->first;
        const size_t rest = it->second - n;
        free_.erase(it);
        if (rest) free_[off + n] = rest;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t n) {
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  int type = 0;
  std::vector<uint64_t> reqs;
  size_t off = 0, bytes = 0;
  int paths = 0;
  // pinned host staging
  hcx_complex* h_params = nullptr;
  hcx_complex* h_roots = nullptr;
  uint32_t* h_flags = nullptr;
  size_t cap_params = 0, cap_roots = 0, cap_paths = 0;
  // the slot's launch workspace (path queue, output descriptor, instance
  // counters, FP64 re-track list): one launch at a time uses it
  hcx::WsBlock ws;
  unsigned long long* d_inst_done = nullptr;
};

// (Re)allocates a slot's workspace block; only while the slot is idle.
bool slot_ws_alloc(Slot& sl, long long inst, long long paths) {
  if (sl.ws.base) (void)hipFree(sl.ws.base);
  sl.ws = hcx::WsBlock();
  void* p = nullptr;
  if (hipMalloc(&p, hcx::ws_layout(sizeof(hcx::SegOut), inst, paths).bytes) != hipSuccess) return false;
  sl.ws.base = p;
  sl.ws.inst_cap = inst;
  sl.ws.paths_cap = paths;
  return true;
}

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

}  // namespace

struct hcx_server {
  hcx_server_opts so{};
  hcx_track_opts to{};
  std::vector<hcx_problem*> types;
  std::mutex mu;
  std::condition_variable cv;
  std::map<uint64_t, Request> reqs;
  std::vector<std::deque<uint64_t>> queues;
  uint64_t next_id = 1;
  std::vector<Slot> slots;
  Arena arena;
  char* d_arena = nullptr;
  std::thread worker;
  bool stop = false;
  bool started = false;
  int rr = 0;
  hcx_server_stats st{};
  std::vector<double> lat_ms;
  std::string error;

  size_t request_bytes(const hcx_problem* p) const {
    return align256((size_t)p->k * sizeof(hcx_complex)) +
           align256((size_t)(p->C + 1) * (2 * p->M + 1) * sizeof(hcx_complex)) +
           align256((size_t)p->K * p->n * sizeof(hcx_complex)) + align256((size_t)p->K * sizeof(uint32_t));
  }

  // Results of one finished launch, unpacked outside the mutex.
  struct Done {
    Slot* slot;
    std::vector<std::vector<hcx_complex>> roots;
    std::vector<std::vector<uint32_t>> flags;
    std::string err;
  };

  // Collects finished launches without the mutex (scheduler thread only).
  void poll_slots(std::vector<Done>& out) {
    for (Slot& s : slots) {
      if (!s.busy) continue;
      const hipError_t q = hipEventQuery(s.done);
      if (q == hipErrorNotReady) continue;
      Done d;
      d.slot = &s;
      if (q != hipSuccess) d.err = hipGetErrorString(q);
      const hcx_problem* p = types[s.type];
      const size_t nr = (size_t)p->K * p->n, nf = (size_t)p->K;
      d.roots.resize(s.reqs.size());
      d.flags.resize(s.reqs.size());
      for (size_t i = 0; i < s.reqs.size(); ++i) {
        d.roots[i].assign(s.h_roots + i * nr, s.h_roots + (i + 1) * nr);
        d.flags[i].assign(s.h_flags + i * nf, s.h_flags + (i + 1) * nf);
      }
      out.push_back(std::move(d));
    }
  }

  // Publishes finished launches (mutex held).
  void reap(std::vector<Done>& done) {
    const auto now = Clock::now();
    for (Done& d : done) {
      Slot& s = *d.slot;
      if (!d.err.empty() && error.empty()) error = d.err;
      const hcx_problem* p = types[s.type];
      for (size_t i = 0; i < s.reqs.size(); ++i) {
        auto it = reqs.find(s.reqs[i]);
        if (it == reqs.end()) continue;
        Request& r = it->second;
        r.roots = std::move(d.roots[i]);
        r.flags = std::move(d.flags[i]);
        r.refined = r.real = 0;
        for (uint32_t f : r.flags) {
          r.refined += (f & HCX_FLAG_REFINED) ? 1 : 0;
          r.real += (f & HCX_FLAG_REAL) ? 1 : 0;
        }
        r.t_done = now;
        r.done = true;
        lat_ms.push_back(std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count());
        st.completed++;
        st.paths_done += p->K;
        st.converged_paths += r.refined;
      }
      st.inflight_requests -= (int64_t)s.reqs.size();
      st.inflight_paths -= s.paths;
      st.hbm_in_use -= (int64_t)s.bytes;
      arena.release(s.off, s.bytes);
      s.reqs.clear();
      s.busy = false;
    }
    done.clear();
  }

  hcx_status ensure_staging(Slot& s, size_t params, size_t roots, size_t paths) {
    if (s.cap_params < params) {
      if (s.h_params) (void)hipHostFree(s.h_params);
      if (hipHostMalloc((void**)&s.h_params, params * sizeof(hcx_complex), hipHostMallocDefault) != hipSuccess)
        return HCX_ERR_OUT_OF_MEMORY;
      s.cap_params = params;
    }
    if (s.cap_roots < roots) {
      if (s.h_roots) (void)hipHostFree(s.h_roots);
      if (hipHostMalloc((void**)&s.h_roots, roots * sizeof(hcx_complex), hipHostMallocDefault) != hipSuccess)
        return HCX_ERR_OUT_OF_MEMORY;
      s.cap_roots = roots;
    }
    if (s.cap_paths < paths) {
      if (s.h_flags) (void)hipHostFree(s.h_flags);
      if (hipHostMalloc((void**)&s.h_flags, paths * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
        return HCX_ERR_OUT_OF_MEMORY;
      s.cap_paths = paths;
    }
    return HCX_OK;
  }

  // Picks one micro-batch for a free slot if any type has queued requests
  // that fit the budget (mutex held): pops the requests, reserves the arena
  // block and stages the parameters in pinned memory.  Returns the slot or
  // nullptr.  The HIP work is issued by issue() without the mutex.
  Slot* pick() {
    Slot* slot = nullptr;
    for (Slot& s : slots)
      if (!s.busy) {
        slot = &s;
        break;
      }
    if (!slot) return nullptr;
    const int nt = (int)types.size();
    for (int q = 0; q < nt; ++q) {
      const int t = (rr + q) % nt;
      if (queues[t].empty()) continue;
      hcx_problem* p = types[t];
      const size_t per = request_bytes(p);
      int B = std::min<int>((int)queues[t].size(), so.max_batch_requests);
      size_t off = 0;
      while (B > 0 && !arena.alloc(per * (size_t)B, off)) B /= 2;
      if (B == 0) continue;  // budget full for this type; try others
      rr = (t + 1) % nt;
      Slot& s = *slot;
      if (ensure_staging(s, (size_t)B * p->k, (size_t)B * p->K * p->n, (size_t)B * p->K) != HCX_OK) {
        arena.release(off, per * (size_t)B);
        if (error.empty()) error = "pinned staging allocation failed";
        return nullptr;
      }
      s.type = t;
      s.off = off;
      s.bytes = per * (size_t)B;
      s.paths = B * p->K;
      s.reqs.clear();
      const auto now = Clock::now();
      for (int i = 0; i < B; ++i) {
        const uint64_t id = queues[t].front();
        queues[t].pop_front();
        Request& r = reqs[id];
        std::memcpy(s.h_params + (size_t)i * p->k, r.params.data(), sizeof(hcx_complex) * p->k);
        r.t_launch = now;
        s.reqs.push_back(id);
      }
      s.busy = true;
      st.launches++;
      st.inflight_requests += B;
      st.inflight_paths += s.paths;
      st.peak_inflight_paths = std::max(st.peak_inflight_paths, st.inflight_paths);
      st.hbm_in_use += (int64_t)s.bytes;
      return slot;
    }
    return nullptr;
  }

  // Issues the copies, the launch and the completion event of a picked slot
  // (no mutex: only the scheduler thread touches slots).
  void issue(Slot& s) {
    hcx_problem* p = types[s.type];
    const int B = (int)s.reqs.size();
    char* base = d_arena + s.off;
    hcx_complex* d_params = reinterpret_cast<hcx_complex*>(base);
    base += align256((size_t)p->k * sizeof(hcx_complex)) * B;
    hcx_complex* d_work = reinterpret_cast<hcx_complex*>(base);
    base += align256((size_t)(p->C + 1) * (2 * p->M + 1) * sizeof(hcx_complex)) * B;
    hcx_complex* d_roots = reinterpret_cast<hcx_complex*>(base);
    base += align256((size_t)p->K * p->n * sizeof(hcx_complex)) * B;
    uint32_t* d_flags = reinterpret_cast<uint32_t*>(base);
    (void)hipMemcpyAsync(d_params, s.h_params, sizeof(hcx_complex) * (size_t)B * p->k, hipMemcpyHostToDevice,
                         s.stream);
    hcx_batch_desc b;
    std::memset(&b, 0, sizeof(b));
    b.num_instances = B;
    b.d_target_params = d_params;
    b.d_coeff_work = d_work;
    hcx_result_desc r;
    std::memset(&r, 0, sizeof(r));
    r.d_roots = d_roots;
    r.d_flags = d_flags;
    r.d_inst_done = reinterpret_cast<uint64_t*>(s.d_inst_done);
    const hcx::LaunchWork w = hcx::seg_work(s.ws, 0, 0, 0);
    hcx_problem* pp = p;
    const hcx_status ls = hcx::track_impl(1, &pp, &b, &r, &to, s.stream, 1, nullptr, &w);
    if (ls != HCX_OK) {
      std::lock_guard<std::mutex> lk(mu);
      if (error.empty()) error = hcx_last_error();
    }
    (void)hipMemcpyAsync(s.h_roots, d_roots, sizeof(hcx_complex) * (size_t)B * p->K * p->n, hipMemcpyDeviceToHost,
                         s.stream);
    (void)hipMemcpyAsync(s.h_flags, d_flags, sizeof(uint32_t) * (size_t)B * p->K, hipMemcpyDeviceToHost, s.stream);
    (void)hipEventRecord(s.done, s.stream);
  }

  bool idle() const {
    for (const auto& q : queues)
      if (!q.empty()) return false;
    for (const Slot& s : slots)
      if (s.busy) return false;
    return true;
  }

  void run() {
    std::vector<Done> done;
    std::vector<Slot*> picked;
    while (true) {
      poll_slots(done);
      {
        std::unique_lock<std::mutex> lk(mu);
        reap(done);
        picked.clear();
        while (Slot* s = pick()) picked.push_back(s);
        bool busy = false;
        for (const Slot& s : slots) busy = busy || s.busy;
        if (picked.empty()) {
          if (stop && idle()) break;
          cv.notify_all();
          // wait for a submission, or poll the in-flight launches every 50 us
          cv.wait_for(lk, std::chrono::microseconds(busy ? 50 : 1000));
          continue;
        }
      }
      for (Slot* s : picked) issue(*s);
      cv.notify_all();
    }
    cv.notify_all();
  }
};

extern "C" {

void hcx_server_opts_default(hcx_server_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->hbm_budget_bytes = (int64_t)1 << 30;
  o->max_batch_requests = 64;
  o->streams = 4;
  o->max_queue = 1 << 20;
}

hcx_status hcx_server_create(const hcx_server_opts* so, const hcx_track_opts* to, hcx_server** out) {
  hcx::clear_error();
  if (!so || !to || !out || so->hbm_budget_bytes < (1 << 20) || so->max_batch_requests < 1 || so->streams < 1 ||
      so->streams > 4 || so->max_queue < 1) {
    hcx::set_error("hcx_server_create: invalid options (budget >= 1 MiB, batch >= 1, 1..4 streams)");
    return HCX_ERR_INVALID_ARG;
  }
  if (hcx_device_count() < 1) {
    hcx::set_error("hcx_server_create: no HIP device available");
    return HCX_ERR_NO_DEVICE;
  }
  auto* s = new hcx_server();
  s->so = *so;
  s->to = *to;
  // no replica groups: a request's launch shape (alone or in a mix of
  // types) then never changes its term sums, so its results are the same
  // bit for bit whatever else is on the GPU
  s->to.replicas = 1;
  hipError_t e = hipMalloc(&s->d_arena, (size_t)so->hbm_budget_bytes);
  if (e != hipSuccess) {
    hcx::set_error("hcx_server_create: cannot allocate the %lld-byte budget: %s", (long long)so->hbm_budget_bytes,
                   hipGetErrorString(e));
    delete s;
    return HCX_ERR_OUT_OF_MEMORY;
  }
  s->arena.init(s->d_arena, (size_t)so->hbm_budget_bytes);
  s->slots.resize(so->streams);
  for (Slot& sl : s->slots) {
    bool ok = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) == hipSuccess &&
              slot_ws_alloc(sl, so->max_batch_requests, 0) &&
              hipMalloc(&sl.d_inst_done, sizeof(unsigned long long) * so->max_batch_requests) == hipSuccess;
    if (!ok) {
      hcx::set_error("hcx_server_create: stream/workspace allocation failed");
      hcx_server_destroy(s);
      return HCX_ERR_HIP;
    }
  }
  *out = s;
  return HCX_OK;
}

hcx_status hcx_server_add_type(hcx_server* s, hcx_problem* p, int* type_id) {
  hcx::clear_error();
  if (!s || !p || !type_id) {
    hcx::set_error("hcx_server_add_type: null argument");
    return HCX_ERR_INVALID_ARG;
  }
  if (p->k < 1 || (!p->d_cmap_scalar && (p->M != 1 || !p->d_affine))) {
    hcx::set_error("hcx_server_add_type: the problem needs a parameter coefficient map (affine or polynomial)");
    return HCX_ERR_INVALID_ARG;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->started) {
    hcx::set_error("hcx_server_add_type: types must be added before the first submit");
    return HCX_ERR_INVALID_ARG;
  }
  if (s->request_bytes(p) > (size_t)s->so.hbm_budget_bytes) {
    hcx::set_error("hcx_server_add_type: one request of this type exceeds the HBM budget");
    return HCX_ERR_BUDGET;
  }
  // re-track path lists hold every path of a full batch of the largest type
  // (no launch is in flight before the first submit)
  const long long need = (long long)s->so.max_batch_requests * p->K;
  for (Slot& sl : s->slots) {
    if (sl.ws.paths_cap >= need) continue;
    if (!slot_ws_alloc(sl, s->so.max_batch_requests, need)) {
      hcx::set_error("hcx_server_add_type: cannot allocate the launch workspace");
      return HCX_ERR_OUT_OF_MEMORY;
    }
  }
  s->types.push_back(p);
  s->queues.emplace_back();
  *type_id = (int)s->types.size() - 1;
  return HCX_OK;
}

hcx_status hcx_server_submit(hcx_server* s, int type_id, const hcx_complex* h_target_params, uint64_t* request_id) {
  hcx::clear_error();
  if (!s || !h_target_params || !request_id) {
    hcx::set_error("hcx_server_submit: null argument");
    return HCX_ERR_INVALID_ARG;
  }
  std::unique_lock<std::mutex> lk(s->mu);
  if (type_id < 0 || type_id >= (int)s->types.size()) {
    hcx::set_error("hcx_server_submit: unknown type %d", type_id);
    return HCX_ERR_INVALID_ARG;
  }
  if (!s->error.empty()) {
    hcx::set_error("hcx_server_submit: server failed: %s", s->error.c_str());
    return HCX_ERR_HIP;
  }
  size_t queued = 0;
  for (auto& q : s->queues) queued += q.size();
  if ((int64_t)queued >= s->so.max_queue) {
    s->st.rejected++;
    hcx::set_error("hcx_server_submit: queue full (%zu)", queued);
    return HCX_ERR_BUDGET;
  }
  if (!s->started) {
    s->started = true;
    s->worker = std::thread([s] { s->run(); });
  }
  const uint64_t id = s->next_id++;
  Request& r = s->reqs[id];
  r.type = type_id;
  r.params.assign(h_target_params, h_target_params + s->types[type_id]->k);
  r.t_submit = Clock::now();
  s->queues[type_id].push_back(id);
  s->st.submitted++;
  *request_id = id;
  lk.unlock();
  s->cv.notify_all();
  return HCX_OK;
}

hcx_status hcx_server_poll(hcx_server* s, uint64_t request_id, int* done, hcx_request_result* out,
                           hcx_complex* h_roots, uint32_t* h_flags) {
  hcx::clear_error();
  if (!s || !done) {
    hcx::set_error("hcx_server_poll: null argument");
    return HCX_ERR_INVALID_ARG;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(request_id);
  if (it == s->reqs.end()) {
    hcx::set_error("hcx_server_poll: unknown or already collected request %llu", (unsigned long long)request_id);
    return HCX_ERR_INVALID_ARG;
  }
  Request& r = it->second;
  *done = r.done ? 1 : 0;
  if (!r.done) return HCX_OK;
  if (!out && !h_roots && !h_flags) return HCX_OK;  // status query only
  const hcx_problem* p = s->types[r.type];
  if (out) {
    out->type_id = r.type;
    out->num_paths = p->K;
    out->refined = r.refined;
    out->real = r.real;
    out->latency_ms = std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count();
    out->queue_ms = std::chrono::duration<double, std::milli>(r.t_launch - r.t_submit).count();
  }
  if (h_roots) std::memcpy(h_roots, r.roots.data(), r.roots.size() * sizeof(hcx_complex));
  if (h_flags) std::memcpy(h_flags, r.flags.data(), r.flags.size() * sizeof(uint32_t));
  s->reqs.erase(it);
  return HCX_OK;
}

hcx_status hcx_server_drain(hcx_server* s, double timeout_s) {
  hcx::clear_error();
  if (!s) {
    hcx::set_error("hcx_server_drain: null server");
    return HCX_ERR_INVALID_ARG;
  }
  std::unique_lock<std::mutex> lk(s->mu);
  const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(timeout_s));
  while (!s->idle()) {
    if (!s->started) break;
    if (s->cv.wait_until(lk, deadline) == std::cv_status::timeout && !s->idle()) {
      hcx::set_error("hcx_server_drain: timeout with work in flight");
      return HCX_ERR_NOT_READY;
    }
  }
  if (!s->error.empty()) {
    hcx::set_error("hcx_server_drain: %s", s->error.c_str());
    return HCX_ERR_HIP;
  }
  return HCX_OK;
}

hcx_status hcx_server_get_stats(hcx_server* s, hcx_server_stats* out) {
  hcx::clear_error();
  if (!s || !out) {
    hcx::set_error("hcx_server_get_stats: null argument");
    return HCX_ERR_INVALID_ARG;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  std::vector<double> v = s->lat_ms;
  if (!v.empty()) {
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {
      const double pos = q * (v.size() - 1);
      const size_t i = (size_t)pos;
      const double f = pos - i;
      return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
    };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    double sum = 0;
    for (double x : v) sum += x;
    out->mean_ms = sum / v.size();
    out->max_ms = v.back();
  }
  return HCX_OK;
}

hcx_status hcx_server_reset_stats(hcx_server* s) {
  if (!s) return HCX_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  const int64_t inflight_r = s->st.inflight_requests, inflight_p = s->st.inflight_paths, hbm = s->st.hbm_in_use;
  s->st = hcx_server_stats{};
  s->st.inflight_requests = inflight_r;
  s->st.inflight_paths = inflight_p;
  s->st.hbm_in_use = hbm;
  s->lat_ms.clear();
  return HCX_OK;
}

hcx_status hcx_server_destroy(hcx_server* s) {
  if (!s) return HCX_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (Slot& sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.h_params) (void)hipHostFree(sl.h_params);
    if (sl.h_roots) (void)hipHostFree(sl.h_roots);
    if (sl.h_flags) (void)hipHostFree(sl.h_flags);
    if (sl.ws.base) (void)hipFree(sl.ws.base);
    if (sl.d_inst_done) (void)hipFree(sl.d_inst_done);
  }
  if (s->d_arena) (void)hipFree(s->d_arena);
  delete s;
  return HCX_OK;
}

}  // extern "C"
