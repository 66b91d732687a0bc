// Randomised check of Poisson::mgpx::Arena (poisson/arena.h) on a host
// buffer: blocks are aligned, inside the region and never overlap; a take
// that fails leaves no free range large enough; giving every block back
// coalesces the region into one free range; foreign pointers are refused;
// best fit picks the smallest adequate range.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "arena.h"

#define REQUIRE(c)                                                   \
    do {                                                             \
        if (!(c)) {                                                  \
            std::fprintf(stderr, "FAILED %s line %d\n", #c, __LINE__); \
            std::exit(1);                                            \
        }                                                            \
    } while (0)

using Poisson::mgpx::Arena;

int main() {
    const size_t size = (size_t)64 << 20;
    std::vector<char> mem(size + Arena::ALIGN);
    char* base = mem.data();
    Arena a(base, size);
    REQUIRE(a.size() == size && a.free_ranges() == 1 && a.in_use() == 0);
    std::mt19937_64 rng(7);
    struct Blk { char* p; size_t n; };
    std::vector<Blk> live;
    size_t fails = 0;
    for (int step = 0; step < 200000; step++) {
        if (live.empty() || rng() % 100 < 55) {
            const size_t n = 1 + rng() % (rng() % 10 == 0 ? (4u << 20) : 20000u);
            char* p = static_cast<char*>(a.take(n));
            if (!p) {
                fails++;
                continue;
            }
            REQUIRE((size_t)(p - base) % Arena::ALIGN == 0);
            REQUIRE(p >= base && p + n <= base + size);
            for (const Blk& b : live) REQUIRE(p + n <= b.p || b.p + b.n <= p);
            live.push_back({p, n});
        } else {
            const size_t k = rng() % live.size();
            REQUIRE(a.give(live[k].p));
            REQUIRE(!a.give(live[k].p));  // twice: refused
            live[k] = live.back();
            live.pop_back();
        }
    }
    REQUIRE(fails > 0);  // the region filled up at times
    size_t used = 0;
    for (const Blk& b : live) used += (b.n + Arena::ALIGN - 1) / Arena::ALIGN * Arena::ALIGN;
    REQUIRE(a.in_use() == used);
    for (const Blk& b : live) REQUIRE(a.give(b.p));
    REQUIRE(a.free_ranges() == 1 && a.in_use() == 0);
    REQUIRE(a.take(size) == base);  // the whole region again
    REQUIRE(a.take(1) == nullptr);
    REQUIRE(a.give(base));
    int foreign = 0;
    REQUIRE(!a.give(&foreign) || !a.owns(&foreign));
    REQUIRE(!a.owns(base + size));
    // best fit: free ranges of 3, 1 and 2 units; a 1-unit take gets the 1-unit hole
    const size_t U = Arena::ALIGN;
    char* p0 = (char*)a.take(3 * U); char* s0 = (char*)a.take(U);
    char* p1 = (char*)a.take(U);     char* s1 = (char*)a.take(U);
    char* p2 = (char*)a.take(2 * U); char* s2 = (char*)a.take(U);
    REQUIRE(p0 && s0 && p1 && s1 && p2 && s2);
    char* rest = (char*)a.take(size - 9 * U);
    REQUIRE(rest && a.take(1) == nullptr);
    REQUIRE(a.give(p0) && a.give(p1) && a.give(p2));
    REQUIRE(a.free_ranges() == 3);
    REQUIRE(a.take(U) == p1);
    REQUIRE(a.take(2 * U) == p2);
    REQUIRE(a.take(2 * U) == p0);
    std::printf("arena ok (%zu failed takes)\n", fails);
    return 0;
}
