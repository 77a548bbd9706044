// Launch sequences replayed as HIP graphs.
//
// The offset loops of the overlap and exact-matching stages issue hundreds
// of small kernels whose grids depend only on host-known bounds; their
// data-dependent sizes are read from device counters inside the kernels.
// Such a loop is a pure function of (its parameters, the buffers it
// touches), so it can be captured once and replayed.  A key names the
// parameters and the caller's buffers; the context's scratch arena is
// covered by its reallocation counter (Scratch::gen): an entry is valid
// only while no scratch buffer has moved since it was made.
//
// Policy per key: the first sighting runs eagerly (growing whatever
// scratch it needs); a later sighting with the arena unchanged captures
// the loop on the stream, instantiates it and launches it; further
// sightings replay.  A capture that fails (for instance because the loop
// had to allocate) falls back to running the same launches eagerly, and
// the key is not captured again.  Enqueue functions must only enqueue
// device work on the stream (no host side effects), since a failed capture
// enqueues them twice.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>

#include "prims.cuh"

namespace curc {

struct GraphKey {
    static constexpr int N = 20;
    uint64_t v[N] = {};
    bool operator<(const GraphKey& o) const { return std::lexicographical_compare(v, v + N, o.v, o.v + N); }
};

struct GraphCache {
    struct Entry {
        hipGraphExec_t exec = nullptr;
        uint64_t gen = ~0ull;
        bool bad = false;
        uint64_t last_use = 0;
    };
    static constexpr size_t MAX_ENTRIES = 48;
    std::map<GraphKey, Entry> m;
    bool enabled = true;
    uint64_t tick = 0, captures = 0, replays = 0, eager = 0;

    template <class F>
    void run(const GraphKey& k, Scratch& S, hipStream_t st, F&& enqueue) {
        if (!enabled) {
            enqueue();
            return;
        }
        ++tick;
        auto it = m.find(k);
        if (it == m.end()) {
            evict_if_full();
            enqueue();
            ++eager;
            Entry e;
            e.gen = S.gen;
            e.last_use = tick;
            m.emplace(k, e);
            return;
        }
        Entry& e = it->second;
        e.last_use = tick;
        if (e.exec && e.gen == S.gen) {
            CURC_HIP(hipGraphLaunch(e.exec, st));
            ++replays;
            return;
        }
        if (e.exec) {
            hipGraphExecDestroy(e.exec);
            e.exec = nullptr;
        }
        if (e.bad || e.gen != S.gen) {
            enqueue();
            ++eager;
            e.gen = S.gen;
            return;
        }
        // the arena has not moved since the last eager run of this key
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        bool ok = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            try {
                enqueue();
            } catch (...) {
                ok = false;
            }
            // always leave capture mode, whatever enqueue did
            const bool ended = hipStreamEndCapture(st, &g) == hipSuccess;
            ok = ok && ended && g != nullptr && S.gen == e.gen;
        }
        if (ok) ok = hipGraphInstantiate(&x, g, nullptr, nullptr, 0) == hipSuccess;
        if (g) hipGraphDestroy(g);
        if (!ok) {
            (void)hipGetLastError();
            if (x) hipGraphExecDestroy(x);
            e.bad = true;
            enqueue();
            ++eager;
            e.gen = S.gen;
            return;
        }
        e.exec = x;
        ++captures;
        CURC_HIP(hipGraphLaunch(e.exec, st));
    }

    void evict_if_full() {
        if (m.size() < MAX_ENTRIES) return;
        auto lru = m.begin();
        for (auto it = m.begin(); it != m.end(); ++it)
            if (it->second.last_use < lru->second.last_use) lru = it;
        if (lru->second.exec) hipGraphExecDestroy(lru->second.exec);
        m.erase(lru);
    }

    void clear() {
        for (auto& kv : m)
            if (kv.second.exec) hipGraphExecDestroy(kv.second.exec);
        m.clear();
    }
};

}  // namespace curc
