// Memory regions carved on the host, for callers that build and drop many
// hierarchies in one process (the serving loop).  hipMalloc per array costs
// tens of microseconds and hipFree / hipHostFree wait for the whole device,
// every stream included, so a server that allocated each new shape's arrays
// and freed evicted ones stalled all running solves at every eviction.  An
// Arena is one allocation made up front; blocks are taken from it best-fit
// and given back coalesced, with no HIP call.  A block given back must not be
// in use by queued work: the server returns a hierarchy's blocks only after
// its stream has drained.
//
// The Poisson classes allocate through dev_alloc / host_alloc: from the
// calling thread's arena when one is installed (ArenaScope), else with
// mgp_malloc / mgp_host_alloc, to which they also fall back when the arena
// has no free range large enough.  dev_free / host_free find the owning arena
// from the address (registered arenas), so a block may be freed from any
// thread.  No counterpart in the reference (one solve per process).
#ifndef POISSON_ARENA
#define POISSON_ARENA

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "mgp.h"

namespace Poisson {
namespace mgpx {

class Arena {
public:
    static constexpr size_t ALIGN = 256;

    Arena(void* base, size_t bytes) : base_(static_cast<char*>(base)), size_((bytes / ALIGN) * ALIGN) {
        if (size_) free_[0] = size_;
    }
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;

    char* base() const { return base_; }
    size_t size() const { return size_; }
    bool owns(const void* p) const {
        const char* c = static_cast<const char*>(p);
        return c >= base_ && c < base_ + size_;
    }

    // the smallest free range that fits (ties: lowest offset); nullptr if none
    void* take(size_t bytes) {
        const size_t need = (bytes + ALIGN - 1) / ALIGN * ALIGN;
        if (need == 0) return nullptr;
        std::lock_guard<std::mutex> lk(mu_);
        auto best = free_.end();
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= need && (best == free_.end() || it->second < best->second)) best = it;
        if (best == free_.end()) return nullptr;
        const size_t off = best->first, len = best->second;
        free_.erase(best);
        if (len > need) free_[off + need] = len - need;
        used_[off] = need;
        in_use_ += need;
        return base_ + off;
    }

    // back into the free list, merged with free neighbours; false if p is not
    // a block of this arena
    bool give(void* p) {
        std::lock_guard<std::mutex> lk(mu_);
        const size_t off = (size_t)(static_cast<char*>(p) - base_);
        auto u = used_.find(off);
        if (u == used_.end()) return false;
        size_t start = off, len = u->second;
        used_.erase(u);
        in_use_ -= len;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && next->first == off + len) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = len;
        return true;
    }

    size_t in_use() {
        std::lock_guard<std::mutex> lk(mu_);
        return in_use_;
    }
    size_t free_ranges() {
        std::lock_guard<std::mutex> lk(mu_);
        return free_.size();
    }

private:
    std::mutex mu_;
    char* base_;
    size_t size_;
    size_t in_use_ = 0;
    std::map<size_t, size_t> free_;            // offset -> length, no two adjacent
    std::unordered_map<size_t, size_t> used_;  // offset -> length
};

// arenas that frees are matched against (an arena is registered for its
// whole life by its owner)
struct ArenaRegistry {
    std::mutex mu;
    std::vector<Arena*> list;
};
inline ArenaRegistry& arena_registry() {
    static ArenaRegistry r;
    return r;
}
inline void register_arena(Arena* a) {
    std::lock_guard<std::mutex> lk(arena_registry().mu);
    arena_registry().list.push_back(a);
}
inline void unregister_arena(Arena* a) {
    std::lock_guard<std::mutex> lk(arena_registry().mu);
    auto& v = arena_registry().list;
    for (size_t i = 0; i < v.size(); i++)
        if (v[i] == a) {
            v.erase(v.begin() + (std::ptrdiff_t)i);
            return;
        }
}
inline bool give_to_arena(void* p) {
    std::lock_guard<std::mutex> lk(arena_registry().mu);
    for (Arena* a : arena_registry().list)
        if (a->owns(p)) return a->give(p);
    return false;
}

inline Arena*& device_arena_slot() {
    thread_local Arena* a = nullptr;
    return a;
}
inline Arena*& host_arena_slot() {
    thread_local Arena* a = nullptr;
    return a;
}

// the calling thread's arenas (device, pinned host) until the scope ends;
// nullptr for direct allocation
struct ArenaScope {
    Arena *prev_dev, *prev_host;
    ArenaScope(Arena* dev, Arena* host) : prev_dev(device_arena_slot()), prev_host(host_arena_slot()) {
        device_arena_slot() = dev;
        host_arena_slot() = host;
    }
    ~ArenaScope() {
        device_arena_slot() = prev_dev;
        host_arena_slot() = prev_host;
    }
    ArenaScope(const ArenaScope&) = delete;
    ArenaScope& operator=(const ArenaScope&) = delete;
};

// device memory: the thread's arena, else mgp_malloc (MGP_OK or its code)
inline int dev_alloc(void** p, size_t bytes) {
    if (Arena* a = device_arena_slot())
        if ((*p = a->take(bytes))) return MGP_OK;
    return mgp_malloc(p, bytes);
}
inline void dev_free(void* p) {
    if (!p) return;
    if (!give_to_arena(p)) mgp_free(p);
}
// pinned host memory, likewise
inline int host_alloc(void** p, size_t bytes) {
    if (Arena* a = host_arena_slot())
        if ((*p = a->take(bytes))) return MGP_OK;
    return mgp_host_alloc(p, bytes);
}
inline void host_free(void* p) {
    if (!p) return;
    if (!give_to_arena(p)) mgp_host_free(p);
}

}  // namespace mgpx
}  // namespace Poisson

#endif
