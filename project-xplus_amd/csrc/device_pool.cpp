// Caching device allocator behind every DevBuf of the library.
//
// hipMalloc / hipFree of buffers of a few MB cost 0.1-0.25 ms each on
// MI355X (a solve's set-up allocates dozens of them: SELL arrays, solver
// vectors, the multilevel hierarchy), while blocks below 256 KB come from the
// runtime's own cache in under a microsecond.  Blocks of at least kMinCached
// bytes are therefore kept after release and handed out again: sizes are
// rounded up to one of eight classes per octave (at most 12.5% slack), a
// released block first waits in a pending list (kernels queued before the
// release may still read it), and the first request that wants a pending
// block synchronises the device once and makes every pending block
// reusable.  A failing hipMalloc returns every cached block to the runtime
// and tries again.  XPLUS_NO_POOL=1 in the environment turns caching off.
#include <cstdlib>
#include <mutex>
#include <unordered_map>
#include <utility>
#include <vector>

#include "hip_util.hpp"

namespace xp {
namespace {

constexpr size_t kMinCached = size_t{1} << 18;

size_t size_class(size_t bytes) {
    size_t top = kMinCached;
    while (top < bytes) top <<= 1;  // bytes in (top / 2, top]
    const size_t step = top / 16;   // eight classes between top / 2 and top
    return (bytes + step - 1) / step * step;
}

struct Pool {
    std::mutex mu;
    std::unordered_map<size_t, std::vector<void*>> ready;  // class -> blocks free for reuse
    std::vector<std::pair<size_t, void*>> pending;          // released, queued work may still use them
    size_t cached = 0;                                      // bytes in ready + pending
    bool off = false;
};

Pool& pool() {
    static Pool* p = [] {
        auto* q = new Pool;  // never destroyed: DevBufs with static lifetime release into it at exit
        const char* e = std::getenv("XPLUS_NO_POOL");
        q->off = e != nullptr && e[0] == '1';
        return q;
    }();
    return *p;
}

// Makes every pending block reusable (caller holds the lock).
void retire_pending(Pool& P) {
    if (P.pending.empty()) return;
    HIPX(hipDeviceSynchronize());
    for (const auto& b : P.pending) P.ready[b.first].push_back(b.second);
    P.pending.clear();
}

// Returns every cached block to the runtime (caller holds the lock).
void release_cached(Pool& P) {
    retire_pending(P);
    for (auto& kv : P.ready)
        for (void* q : kv.second) (void)hipFree(q);
    P.ready.clear();
    P.cached = 0;
}

}  // namespace

void* pool_alloc(size_t bytes) {
    Pool& P = pool();
    void* p = nullptr;
    if (bytes < kMinCached || P.off) {
        HIPX(hipMalloc(&p, bytes));
        return p;
    }
    const size_t c = size_class(bytes);
    std::lock_guard<std::mutex> lock(P.mu);
    auto take = [&]() -> void* {
        auto it = P.ready.find(c);
        if (it == P.ready.end() || it->second.empty()) return nullptr;
        void* q = it->second.back();
        it->second.pop_back();
        P.cached -= c;
        return q;
    };
    if ((p = take()) != nullptr) return p;
    for (const auto& b : P.pending) {
        if (b.first == c) {
            retire_pending(P);
            if ((p = take()) != nullptr) return p;
            break;
        }
    }
    if (hipMalloc(&p, c) != hipSuccess) {
        (void)hipGetLastError();
        release_cached(P);
        HIPX(hipMalloc(&p, c));
    }
    return p;
}

void pool_free(void* p, size_t bytes) {
    if (p == nullptr) return;
    Pool& P = pool();
    if (bytes < kMinCached || P.off) {
        (void)hipFree(p);
        return;
    }
    std::lock_guard<std::mutex> lock(P.mu);
    const size_t c = size_class(bytes);
    P.pending.emplace_back(c, p);
    P.cached += c;
}

void pool_trim() {
    Pool& P = pool();
    std::lock_guard<std::mutex> lock(P.mu);
    release_cached(P);
}

size_t pool_cached_bytes() {
    Pool& P = pool();
    std::lock_guard<std::mutex> lock(P.mu);
    return P.cached;
}

}  // namespace xp
