// Ownership of launch workspaces (host side).
//
// A tracking launch needs device memory besides its inputs and outputs: the
// per-segment path queue counters, the device copy of the output descriptors,
// per-instance remaining-path counters and the FP64 re-track path list.  The
// reference's launcher needs none of these (its grid is one block per path,
// kernel_HC_Solver_generalized_3views_4pts.cu:299-300, and GPU_HC_Solver
// owns every buffer, GPU_HC_Solver.cpp:104-140).
//
// Rules kept here:
//  * one workspace block per (device, stream): launches on one stream are
//    ordered by the stream, so they share a block; launches on different
//    streams never share one, so one problem serves several streams at once;
//  * a block grows only after the stream has drained the launches that use
//    the old block (the old block is then freed);
//  * a block used while the stream was being captured into a graph is pinned:
//    the graph refers to it for as long as the graph lives, so it is never
//    freed or resized; a later launch that needs more room gets a new block
//    and the pinned one is kept (retired) until the process ends;
//  * under capture a block cannot grow (nothing may wait on the stream or
//    allocate): the launch is refused and the caller is told to run the same
//    launch once uncaptured first.
//
// The pool is a template over the device-memory policy `Dev` so that these
// rules run on the CPU with a stand-in device in the tests:
//   void* Dev::alloc(size_t bytes)      (nullptr on failure)
//   void  Dev::release(void* p)
//   bool  Dev::sync(const void* stream) (waits for the stream's work)
//   bool  Dev::capturing(const void* stream)
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

namespace hcx {

// Fixed part of every workspace block: kSegSlots segments x 64 counter words
// (word 0 the path queue, word 32 the re-track list length), then the
// device copies of the segment output descriptors.
constexpr int kSegSlots = 8;
constexpr size_t kCounterBytes = (size_t)kSegSlots * 64 * sizeof(uint32_t);

constexpr size_t ws_align(size_t n) { return (n + 255) & ~size_t(255); }

// Byte layout of a block holding `inst` instance counters and `paths`
// re-track list entries (both int32) after the fixed part.
struct WsLayout {
  size_t out_off, inst_off, plist_off, bytes;
};
constexpr WsLayout ws_layout(size_t segout_bytes, long long inst, long long paths) {
  WsLayout L{};
  L.out_off = kCounterBytes;
  L.inst_off = ws_align(L.out_off + (size_t)kSegSlots * segout_bytes);
  L.plist_off = ws_align(L.inst_off + sizeof(int32_t) * (size_t)(inst > 0 ? inst : 0));
  L.bytes = ws_align(L.plist_off + sizeof(int32_t) * (size_t)(paths > 0 ? paths : 0));
  return L;
}

struct WsBlock {
  void* base = nullptr;
  long long inst_cap = 0, paths_cap = 0;
  bool pinned = false;  // referenced by a captured graph
};

enum class WsStatus { kOk, kCaptureTooSmall, kOutOfMemory, kSyncFailed };

template <class Dev>
class WorkspacePool {
 public:
  using Key = std::pair<int, const void*>;  // (device, stream)

  struct Entry {
    std::mutex mu;
    WsBlock cur;
    std::vector<WsBlock> retired;  // pinned blocks replaced by larger ones
  };

  // The stream's entry, locked until the lease is dropped: the caller
  // enqueues its launch while holding it, so no other thread can free or
  // replace the block in between.
  struct Lease {
    Entry* entry = nullptr;
    std::unique_lock<std::mutex> lock;
    const WsBlock& block() const { return entry->cur; }
  };

  WorkspacePool(Dev dev, size_t segout_bytes) : dev_(dev), segout_(segout_bytes) {}

  WsStatus acquire(int device, const void* stream, long long inst, long long paths, Lease& out) {
    Entry* e;
    {
      std::lock_guard<std::mutex> g(mu_);
      auto& slot = entries_[Key(device, stream)];
      if (!slot) slot.reset(new Entry());
      e = slot.get();
    }
    std::unique_lock<std::mutex> lk(e->mu);
    const bool cap = dev_.capturing(stream);
    WsBlock& b = e->cur;
    if (!b.base || b.inst_cap < inst || b.paths_cap < paths) {
      if (cap) return WsStatus::kCaptureTooSmall;
      // grow geometrically so that a sequence of growing launches
      // reallocates a few times only
      const long long ni = inst > b.inst_cap ? std::max(inst, 2 * b.inst_cap) : b.inst_cap;
      const long long np = paths > b.paths_cap ? std::max(paths, 2 * b.paths_cap) : b.paths_cap;
      if (b.base) {
        if (b.pinned) {
          e->retired.push_back(b);
        } else {
          if (!dev_.sync(stream)) return WsStatus::kSyncFailed;
          dev_.release(b.base);
        }
        b = WsBlock();
      }
      const WsLayout L = ws_layout(segout_, ni, np);
      void* p = dev_.alloc(L.bytes);
      if (!p) return WsStatus::kOutOfMemory;
      b.base = p;
      b.inst_cap = ni;
      b.paths_cap = np;
      b.pinned = false;
    }
    if (cap) b.pinned = true;
    out.entry = e;
    out.lock = std::move(lk);
    return WsStatus::kOk;
  }

  // Frees the stream's block after its work drained (a pinned block stays).
  WsStatus release_stream(int device, const void* stream) {
    Entry* e = nullptr;
    {
      std::lock_guard<std::mutex> g(mu_);
      auto it = entries_.find(Key(device, stream));
      if (it == entries_.end()) return WsStatus::kOk;
      e = it->second.get();
    }
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->cur.base && !e->cur.pinned) {
      if (!dev_.sync(stream)) return WsStatus::kSyncFailed;
      dev_.release(e->cur.base);
    } else if (e->cur.base) {
      e->retired.push_back(e->cur);
    }
    e->cur = WsBlock();
    return WsStatus::kOk;
  }

  // Introspection for tests.
  size_t retired_count(int device, const void* stream) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = entries_.find(Key(device, stream));
    return it == entries_.end() ? 0 : it->second->retired.size();
  }
  WsBlock current(int device, const void* stream) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = entries_.find(Key(device, stream));
    return it == entries_.end() ? WsBlock() : it->second->cur;
  }

 private:
  Dev dev_;
  size_t segout_;
  std::mutex mu_;
  std::map<Key, std::unique_ptr<Entry>> entries_;
};

}  // namespace hcx
