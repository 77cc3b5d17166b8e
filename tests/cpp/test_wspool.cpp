// CPU test of the launch-workspace ownership rules (csrc/hcx_wspool.hpp)
// with a stand-in device that records every allocation, free and wait.
// Exit status 0 = all checks passed; a failing check prints its line.
#include <cstdio>
#include <chrono>
#include <cstdlib>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "hcx_wspool.hpp"

namespace {

struct Log {
  std::set<void*> live;        // allocated and not yet released
  std::vector<std::string> ev; // "alloc", "free", "sync"
  bool capturing = false;
};

struct FakeDev {
  Log* log;
  void* alloc(size_t n) {
    void* p = std::malloc(n);
    log->live.insert(p);
    log->ev.push_back("alloc");
    return p;
  }
  void release(void* p) {
    if (!log->live.erase(p)) {
      std::printf("double free or foreign pointer\n");
      std::exit(2);
    }
    log->ev.push_back("free");
    std::free(p);
  }
  bool sync(const void*) {
    log->ev.push_back("sync");
    return true;
  }
  bool capturing(const void*) { return log->capturing; }
};

int failures = 0;
#define CHECK(c)                                              \
  do {                                                        \
    if (!(c)) {                                               \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
      ++failures;                                             \
    }                                                         \
  } while (0)

using Pool = hcx::WorkspacePool<FakeDev>;

}  // namespace

int main() {
  Log log;
  Pool pool(FakeDev{&log}, 128);
  const void* s1 = reinterpret_cast<const void*>(0x10);
  const void* s2 = reinterpret_cast<const void*>(0x20);

  // first use allocates; a smaller launch reuses the block
  void* b1;
  {
    Pool::Lease l;
    CHECK(pool.acquire(0, s1, 10, 100, l) == hcx::WsStatus::kOk);
    b1 = l.block().base;
    CHECK(b1 && l.block().inst_cap >= 10 && l.block().paths_cap >= 100);
  }
  {
    Pool::Lease l;
    CHECK(pool.acquire(0, s1, 5, 50, l) == hcx::WsStatus::kOk);
    CHECK(l.block().base == b1);
  }
  CHECK(log.ev.size() == 1);

  // another stream never shares the block
  {
    Pool::Lease l;
    CHECK(pool.acquire(0, s2, 5, 50, l) == hcx::WsStatus::kOk);
    CHECK(l.block().base != b1);
  }
  // the same stream on another device neither
  {
    Pool::Lease l;
    CHECK(pool.acquire(1, s1, 5, 50, l) == hcx::WsStatus::kOk);
    CHECK(l.block().base != b1);
  }

  // growth waits for the stream BEFORE the old block is freed
  log.ev.clear();
  {
    Pool::Lease l;
    CHECK(pool.acquire(0, s1, 1000, 100, l) == hcx::WsStatus::kOk);
    CHECK(l.block().inst_cap >= 1000);
  }
  CHECK(log.ev.size() == 3 && log.ev[0] == "sync" && log.ev[1] == "free" && log.ev[2] == "alloc");
  CHECK(log.live.count(b1) == 0);

  // a block used under capture is pinned: a later, larger uncaptured launch
  // gets a new block and the pinned one is never freed
  void* cap_block;
  log.capturing = true;
  {
    Pool::Lease l;
    CHECK(pool.acquire(0, s1, 500, 50, l) == hcx::WsStatus::kOk);  // fits: no growth
    cap_block = l.block().base;
    CHECK(l.block().pinned);
  }
  log.capturing = false;
  log.ev.clear();
  {
    Pool::Lease l;
    CHECK(pool.acquire(0, s1, 5000, 50, l) == hcx::WsStatus::kOk);
    CHECK(l.block().base != cap_block && !l.block().pinned);
  }
  CHECK(log.live.count(cap_block) == 1);              // still allocated
  CHECK(log.ev.size() == 1 && log.ev[0] == "alloc");   // no free, no wait
  CHECK(pool.retired_count(0, s1) == 1);

  // under capture a block cannot grow: refused, nothing allocated or freed
  log.capturing = true;
  log.ev.clear();
  {
    Pool::Lease l;
    CHECK(pool.acquire(0, s1, 1 << 20, 50, l) == hcx::WsStatus::kCaptureTooSmall);
    CHECK(l.entry == nullptr);
  }
  CHECK(log.ev.empty());
  // first use of a stream under capture is refused too
  {
    Pool::Lease l;
    const void* s3 = reinterpret_cast<const void*>(0x30);
    CHECK(pool.acquire(0, s3, 1, 1, l) == hcx::WsStatus::kCaptureTooSmall);
  }
  log.capturing = false;

  // release frees an unpinned block after waiting, keeps a pinned one
  log.ev.clear();
  void* cur = pool.current(0, s2).base;
  CHECK(pool.release_stream(0, s2) == hcx::WsStatus::kOk);
  CHECK(log.ev.size() == 2 && log.ev[0] == "sync" && log.ev[1] == "free");
  CHECK(log.live.count(cur) == 0 && pool.current(0, s2).base == nullptr);
  log.capturing = true;
  {
    Pool::Lease l;
    CHECK(pool.acquire(0, s1, 1, 1, l) == hcx::WsStatus::kOk);
  }
  log.capturing = false;
  void* pinned = pool.current(0, s1).base;
  CHECK(pool.release_stream(0, s1) == hcx::WsStatus::kOk);
  CHECK(log.live.count(pinned) == 1 && pool.retired_count(0, s1) == 2);

  // a lease holds the stream's entry: a second thread's acquire on the same
  // stream waits until the first lease is dropped
  {
    Pool::Lease l;
    CHECK(pool.acquire(0, s2, 10, 10, l) == hcx::WsStatus::kOk);
    bool done = false;
    std::thread t([&] {
      Pool::Lease l2;
      pool.acquire(0, s2, 10, 10, l2);
      done = true;
    });
    std::this_thread::sleep_for(std::chrono::milliseconds(50));
    CHECK(!done);
    l.lock.unlock();
    t.join();
    CHECK(done);
  }

  // layout: regions are disjoint, 256-byte aligned, in the block
  const hcx::WsLayout L = hcx::ws_layout(200, 37, 1001);
  CHECK(L.out_off >= hcx::kCounterBytes);
  CHECK(L.inst_off >= L.out_off + 8 * 200 && L.inst_off % 256 == 0);
  CHECK(L.plist_off >= L.inst_off + 37 * 4 && L.plist_off % 256 == 0);
  CHECK(L.bytes >= L.plist_off + 1001 * 4);

  if (failures == 0) std::printf("all workspace ownership checks passed\n");
  return failures == 0 ? 0 : 1;
}
