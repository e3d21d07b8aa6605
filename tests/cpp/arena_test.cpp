// Randomised check of rf_arena (host bookkeeping of the serving memory
// budget): live allocations never overlap and stay inside the block, the
// used-byte count matches, and freeing everything coalesces back to one range.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>

#include "../../randomfield_amd/csrc/rf_arena.h"

#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            std::fprintf(stderr, "FAIL %s:%d %s\n", __FILE__, __LINE__, #c); \
            return 1;                                              \
        }                                                          \
    } while (0)

int main() {
    rf_arena a;
    const size_t cap = 1 << 20;
    a.init(cap, 256);
    CHECK(a.capacity() == cap && a.largest_free() == cap && a.free_ranges() == 1);
    std::mt19937_64 rng(7);
    std::map<size_t, size_t> live;   // offset -> rounded size
    size_t used = 0;
    for (int it = 0; it < 200000; it++) {
        const bool doAlloc = live.empty() || (rng() % 100) < 55;
        if (doAlloc) {
            const size_t want = 1 + rng() % 40000;
            size_t off = 0;
            const bool ok = a.alloc(want, &off);
            if (!ok) {
                CHECK(a.largest_free() < a.round(want));
                continue;
            }
            const size_t sz = a.round(want);
            CHECK(off % 256 == 0 && off + sz <= cap);
            auto nx = live.lower_bound(off);
            if (nx != live.end()) CHECK(off + sz <= nx->first);
            if (nx != live.begin()) {
                auto pv = std::prev(nx);
                CHECK(pv->first + pv->second <= off);
            }
            live[off] = sz;
            used += sz;
        } else {
            auto v = live.begin();
            std::advance(v, rng() % live.size());
            CHECK(a.release(v->first));
            used -= v->second;
            live.erase(v);
        }
        CHECK(a.used() == used && a.live() == live.size());
    }
    CHECK(!a.release(12345));           // not an allocation
    for (auto& kv : live) CHECK(a.release(kv.first));
    CHECK(a.used() == 0 && a.free_ranges() == 1 && a.largest_free() == cap);
    // best fit: the smallest hole that fits is taken
    size_t o1, o2, o3, o4, o5;
    CHECK(a.alloc(1024, &o1) && a.alloc(4096, &o2) && a.alloc(512, &o3) && a.alloc(2048, &o4));
    CHECK(a.release(o1) && a.release(o3));   // holes of 1024 and 512
    CHECK(a.alloc(400, &o5) && o5 == o3);
    std::printf("arena ok (peak %zu)\n", a.peak());
    return 0;
}
