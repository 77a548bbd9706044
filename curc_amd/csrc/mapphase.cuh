// Mapping in probe ranges (the learned scorer's checkpoints sit between
// ranges).  Each read's loop state -- best count and position, the cache of
// verified positions and a few work counters -- lives in HBM between
// launches, so a pass split into ranges replays k_map exactly: same probe
// order, same candidate order, same early exit at <= 1 mismatch.  At a
// checkpoint the state yields 32 features per read; in trace mode they are
// recorded with a label (did the rest of the strand pass improve the best
// count?), in learned mode the network's logit decides whether the read's
// remaining probes are tried at all.  Exact verification decides every
// record either way.
#pragma once
#include "kernels.cuh"

namespace curc {

constexpr uint32_t MAP_NFEAT = 32;

// Per-read loop state, structure of arrays over mapping slots.
struct MapState {
    uint64_t* seen;  // [4][n]: the SeenCache slots
    uint32_t* cnt;   // [4][n]: work counters (see map_range_probe)
    uint8_t* stop;   // [n]: 1 when the scorer ended the read's pass
    uint64_t n;      // slots
};

struct MapCounters {
    uint32_t cands = 0;     // verified candidates
    uint32_t votes = 0;     // candidates at the current best position
    uint32_t empty = 0;     // probes with an empty bucket
    uint32_t blast = 0;     // size of the last probe's bucket (capped 12)
    uint32_t second = 255;  // best count that did not win (may be a bound)
    uint32_t bsum = 0;      // summed bucket sizes
    uint32_t near = 0;      // candidates within 2 of the best count
    uint32_t impr = 0;      // improvements of the best count
    uint32_t tbest = 255;   // probe index of the last improvement (255: none)
    uint32_t full = 0;      // probes with a full bucket
    uint32_t gc = 0;        // G/C share of the read, in 1/255

    __device__ void load(const MapState& s, uint64_t i) {
        const uint32_t a = s.cnt[i], b = s.cnt[s.n + i], c = s.cnt[2 * s.n + i], d = s.cnt[3 * s.n + i];
        cands = a;
        votes = b & 0xFF;
        empty = (b >> 8) & 0xFF;
        blast = (b >> 16) & 0xFF;
        second = b >> 24;
        bsum = c & 0xFFFF;
        near = (c >> 16) & 0xFF;
        impr = c >> 24;
        tbest = d & 0xFF;
        full = (d >> 8) & 0xFF;
        gc = (d >> 16) & 0xFF;
    }
    __device__ void store(const MapState& s, uint64_t i) const {
        s.cnt[i] = cands;
        s.cnt[s.n + i] = min(votes, 255u) | min(empty, 255u) << 8 | min(blast, 255u) << 16 | min(second, 255u) << 24;
        s.cnt[2 * s.n + i] = min(bsum, 0xFFFFu) | min(near, 255u) << 16 | min(impr, 255u) << 24;
        s.cnt[3 * s.n + i] = min(tbest, 255u) | min(full, 255u) << 8 | min(gc, 255u) << 16;
    }
};

// G/C bases of a mapping slot's read.
__device__ __forceinline__ uint32_t read_gc(const MapArgs& a, uint64_t i) {
    uint32_t gc = 0;
    if (i < a.n_un) {
        const uint64_t* rd = a.R + (uint64_t)a.un_ids[i] * a.W;
        for (uint32_t k = 0; k < a.W; ++k) {
            const uint32_t nb = min(32u, a.L - 32u * k);
            const uint64_t v = xw(rd, 32u * k, nb);
            // codes 1 (C) and 2 (G): the two bits differ
            gc += (uint32_t)__popcll((v ^ (v >> 1)) & 0x5555555555555555ull);
        }
    } else {
        const uint8_t* ra = a.nascii + (i - a.n_un) * a.L;
        for (uint32_t o = 0; o < a.L; ++o) {
            const uint8_t u = ld8(ra + o) & 0xDF;
            gc += (u == 'C' || u == 'G') ? 1u : 0u;
        }
    }
    return gc;
}

__global__ void k_map_state_reset(MapArgs a, MapState s, const uint32_t* __restrict__ act,
                                  const uint32_t* __restrict__ d_n) {
    GRID_STRIDE(q, *d_n) {
        const uint64_t i = act[q];
        for (int k = 0; k < 4; ++k) s.seen[k * s.n + i] = ~0ull;
        MapCounters c;
        c.gc = (255u * read_gc(a, i) + a.L / 2) / a.L;
        c.store(s, i);
        s.stop[i] = 0;
    }
}

// The 32 features of a read after t_done probes of the pass (t_prev: the
// probe count at the previous checkpoint).
__device__ __forceinline__ void map_features(const MapArgs& a, const MapCounters& c, uint32_t last, bool isN,
                                             bool rc, uint32_t t_done, uint32_t t_prev, float* __restrict__ o) {
    const float np = (float)((a.L - MAP_K) / MAP_READ_STEP + 1);
    const float mx = (float)(a.max_mm + 1);
    float f[MAP_NFEAT];
    const bool none = last > a.max_mm;
    f[0] = (float)t_done / np;
    f[1] = (float)min(last, a.max_mm + 1) / mx;
    f[2] = none ? 0.f : (float)(c.tbest + 1) / np;
    f[3] = (float)c.cands / 64.f;
    f[4] = (float)c.near / 16.f;
    f[5] = (float)c.votes / 16.f;
    f[6] = (float)c.empty / np;
    f[7] = (float)c.bsum / (12.f * (float)max(t_done, 1u));
    f[8] = (float)c.blast / 12.f;
    f[9] = (float)c.gc / 255.f;
    f[10] = rc ? 1.f : 0.f;
    f[11] = isN ? 1.f : 0.f;
    f[12] = (float)a.L / 256.f;
    f[13] = (float)min(c.second, a.max_mm + 1) / mx;
    f[14] = (float)c.impr / 8.f;
    f[15] = none ? 1.f : (float)(t_done - 1 - min(c.tbest, t_done - 1)) / np;
    f[16] = (float)c.full / (float)max(t_done, 1u);
    f[17] = none ? 1.f : 0.f;
#pragma unroll
    for (uint32_t k = 0; k < 7; ++k) f[18 + k] = 0.f;
    if (!none) f[18 + min(max(last, 2u), 8u) - 2] = 1.f;
    f[25] = (!none && c.tbest >= t_prev && c.tbest != 255) ? 1.f : 0.f;
    f[26] = 1.f;
    f[27] = (float)a.max_mm / 64.f;
    f[28] = t_done >= 32 ? 1.f : 0.f;
    f[29] = (float)min(last, 2 * a.max_mm) / (float)a.L;
    f[30] = (float)(a.max_mm + 1 - min(c.second, a.max_mm + 1)) / mx;
    f[31] = 0.f;
#pragma unroll
    for (uint32_t k = 0; k < MAP_NFEAT; ++k) o[k] = f[k];
}

// Bucket statistics of one probe.
__device__ __forceinline__ void range_bucket(MapCounters& c, uint32_t bs) {
    c.blast = bs;
    c.bsum += bs;
    if (bs == MAP_BUCKET) ++c.full;
    if (bs == 0) ++c.empty;
}

// One candidate position mp after its mismatch count cm (exact when below
// `last`, else the partial sum at which verification stopped); returns true
// when the read settles (<= 1 mismatch).
__device__ __forceinline__ bool range_candidate(MapCounters& c, uint32_t& last, uint64_t& w, uint64_t mp,
                                                uint32_t cm, bool isN, bool rc, uint32_t t) {
    if (cm <= last + 2) ++c.near;
    if (cm < last) {
        c.second = last;
        w = (isN ? 1ull : 0ull) | (rc ? 2ull : 0ull) | (mp << 2);
        last = cm;
        c.tbest = t;
        ++c.impr;
        c.votes = 1;
        return last <= MAP_TARGET;
    }
    if (cm < c.second) c.second = cm;
    return false;
}

// A read's loop state between ranges.
struct RangeState {
    SeenCache seen;
    MapCounters c;
    __device__ void load(const MapState& s, uint64_t i) {
        seen.s0 = s.seen[i];
        seen.s1 = s.seen[s.n + i];
        seen.s2 = s.seen[2 * s.n + i];
        seen.s3 = s.seen[3 * s.n + i];
        c.load(s, i);
    }
    __device__ void store(const MapState& s, uint64_t i) const {
        s.seen[i] = seen.s0;
        s.seen[s.n + i] = seen.s1;
        s.seen[2 * s.n + i] = seen.s2;
        s.seen[3 * s.n + i] = seen.s3;
        c.store(s, i);
    }
};

// Probes [t0, t1) of one read, generic form: words through memory, N-reads
// base by base.  Returns with `last` / `w` / st updated.
template <bool RC>
__device__ __forceinline__ void range_read_generic(const MapArgs& a, uint64_t i, uint32_t t0, uint32_t t1,
                                                   uint32_t& last, uint64_t& w, RangeState& st,
                                                   uint32_t& n_probe, uint32_t& n_cand) {
    const bool isN = i >= a.n_un;
    MapCounters& c = st.c;
    // N-reads: their packed rows and masks when given (word by word), else
    // their characters
    const bool nw = isN && a.nmask;
    const uint64_t* rd = !isN ? a.R + (uint64_t)a.un_ids[i] * a.W
                              : (nw ? a.R + (uint64_t)a.nids[i - a.n_un] * a.W : nullptr);
    const uint64_t* mk = nw ? a.nmask + (i - a.n_un) * a.W : nullptr;
    const uint8_t* ra = isN ? a.nascii + (i - a.n_un) * a.L : nullptr;
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t j = MAP_READ_STEP * t;
        if (j + MAP_K > a.L) break;
        ++n_probe;
        const uint64_t km = (isN && !nw) ? kmer_ascii(ra + j, MAP_K) : xw(rd, j, MAP_K);
        const uint32_t h = (uint32_t)(murmur64(km) & a.mask);
        // the compact two-strand index (a.cs) or one strand's slots / CSR lists
        Bk2 b2;
        MapBucket B{0, nullptr};
        if (a.cs) {
            b2 = bk2_fetch<true>(a, h, RC);
            B.n = RC ? b2.nr : b2.nf;
        } else {
            B = map_bucket(a, h);
        }
        range_bucket(c, B.n);
        const uint64_t ls = RC ? (uint64_t)(a.L - MAP_K - j) : (uint64_t)j;
        for (uint32_t p = 0; p < B.n; ++p) {
            const uint32_t pv = a.cs ? bk2_pos<true>(b2, RC, p) : ld32(B.p + p);
            const uint64_t rp = a.start + (uint64_t)pv * MAP_REF_STEP;
            if (rp < ls) continue;
            const uint64_t mp = rp - ls;
            if (mp + a.L > a.pg_len) continue;
            if (last <= a.max_mm && mp == (w >> 2)) ++c.votes;
            if (st.seen.check_insert(mp)) continue;
            ++n_cand;
            ++c.cands;
            uint32_t cm;
            if (!isN) {
                cm = RC ? rc_mm(a.pg, mp, rd, a.L, last - 1) : fwd_mm(a.pg, mp, rd, a.L, last - 1);
            } else if (nw) {
                cm = nread_count<RC>(a, rd, mk, mp, last);
            } else {
                cm = 0;
                for (uint32_t o = 0; o < a.L; ++o) {
                    const uint8_t rb = base_char(base2(a.pg, mp + o));
                    const uint8_t qb = RC ? rev_base_ascii(ra[a.L - 1 - o]) : ra[o];
                    if (rb != qb) {
                        cm++;
                        if (cm >= last) break;
                    }
                }
            }
            if (range_candidate(c, last, w, mp, cm, isN, RC, t)) return;
        }
    }
}

// Probes [t0, t1) of one non-N read with the bucket-inline index (one
// strand's 64-byte slots, or with CS the compact two-strand index) and the
// read in registers (as k_map_fast): the same candidates, counts and
// partial sums as range_read_generic.
template <bool RC, uint32_t WM, bool CS>
__device__ __forceinline__ void range_read_fast(const MapArgs& a, uint64_t i, uint32_t t0, uint32_t t1,
                                                uint32_t& last, uint64_t& w, RangeState& st,
                                                uint32_t& n_probe, uint32_t& n_cand) {
    MapCounters& c = st.c;
    const uint32_t L = a.L, W = a.W;
    const uint32_t np = (L - MAP_K) / MAP_READ_STEP + 1;
    t1 = min(t1, np);
    const uint64_t* rd = a.R + (uint64_t)a.un_ids[i] * W;
    uint64_t r[WM], q[WM];
#pragma unroll
    for (uint32_t k = 0; k < WM; ++k) r[k] = k < W ? ld64(rd + k) : 0ull;
#pragma unroll
    for (uint32_t k = 0; k < WM; ++k) {
        q[k] = 0;
        if (32 * k < L) {
            const uint32_t n = min(32u, L - 32 * k);
            q[k] = RC ? revcomp(xw_pick(r, L - 32 * k - n, n), n) : xw_pick(r, 32 * k, n);
        }
    }
    for (uint32_t tg = t0; tg < t1; tg += MAP_GROUP) {
        uint4 hd[MAP_GROUP];
        Bk2 hb[MAP_GROUP];
        uint32_t hh[MAP_GROUP];
#pragma unroll
        for (uint32_t u = 0; u < MAP_GROUP; ++u) {
            hd[u] = make_uint4(0, 0, 0, 0);
            if constexpr (CS) {
                hb[u].nf = 0;
                hb[u].nr = 0;
                hb[u].nt = 0;
            }
            hh[u] = 0;
            if (tg + u < t1) {
                hh[u] = (uint32_t)(murmur64(xw_pick(r, MAP_READ_STEP * (tg + u), MAP_K)) & a.mask);
                if constexpr (CS)
                    hb[u] = bk2_fetch<true>(a, hh[u], RC);
                else
                    hd[u] = *reinterpret_cast<const uint4*>(a.bk + (uint64_t)hh[u] * MAP_SLOT);
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < MAP_GROUP; ++u) {
            const uint32_t t = tg + u;
            if (t >= t1) break;
            ++n_probe;
            uint32_t nb;
            if constexpr (CS)
                nb = RC ? hb[u].nr : hb[u].nf;
            else
                nb = hd[u].x;
            range_bucket(c, nb);
            const uint32_t j = MAP_READ_STEP * t;
            const uint64_t ls = RC ? (uint64_t)(L - MAP_K - j) : (uint64_t)j;
            for (uint32_t p = 0; p < nb; ++p) {
                uint32_t pv;
                if constexpr (CS)
                    pv = bk2_pos<true>(hb[u], RC, p);
                else
                    pv = p == 0   ? hd[u].y
                         : p == 1 ? hd[u].z
                         : p == 2 ? hd[u].w
                                  : ld32(a.bk + (uint64_t)hh[u] * MAP_SLOT + 1 + p);
                const uint64_t rp = a.start + (uint64_t)pv * MAP_REF_STEP;
                if (rp < ls) continue;
                const uint64_t mp = rp - ls;
                if (mp + L > a.pg_len) continue;
                if (last <= a.max_mm && mp == (w >> 2)) ++c.votes;
                if (st.seen.check_insert(mp)) continue;
                ++n_cand;
                ++c.cands;
                const uint64_t* g0 = a.pg + (mp >> 5);
                uint64_t g[WM + 1];
#pragma unroll
                for (uint32_t k = 0; k <= WM; ++k) g[k] = k <= W ? ld64(g0 + k) : 0ull;
                const uint32_t sh = (uint32_t)(mp & 31u) << 1;
                uint32_t cm = 0;
                const uint32_t bound = last - 1;
#pragma unroll
                for (uint32_t k = 0; k < WM; ++k) {
                    if (32 * k < L && cm <= bound) {
                        const uint32_t n = min(32u, L - 32 * k);
                        const uint64_t win = sh ? (g[k] << sh) | (g[k + 1] >> (64u - sh)) : g[k];
                        cm += mm_words(win >> ((32u - n) << 1), q[k]);
                    }
                }
                if (range_candidate(c, last, w, mp, cm, false, RC, t)) return;
            }
        }
    }
}

// Probes [t0, t1) (probe t reads the k-mer at base 2t) of every read in the
// active list, continuing k_map's loop from the saved state.  WM > 0 takes
// the register form for non-N reads (bucket-inline index, W <= WM; CS: the
// compact two-strand index).
// With x, each read of the list also writes its features after the range
// (row q of x, t1 probes done); reads already settled write a row too.
template <bool RC, uint32_t WM, bool CS>
__global__ void __launch_bounds__(256) k_map_range(MapArgs a, MapState s, const uint32_t* __restrict__ act,
                                                   const uint32_t* __restrict__ d_n, uint32_t t0, uint32_t t1,
                                                   float* __restrict__ x) {
    uint32_t n_probe = 0, n_cand = 0;
    const uint32_t n_act = *d_n;
    GRID_STRIDE(q, n_act) {
        const uint64_t i = act[q];
        uint32_t last = a.last[i];
        const bool isN = i >= a.n_un;
        if (last <= MAP_TARGET) {
            if (x) {
                MapCounters c;
                c.load(s, i);
                map_features(a, c, last, isN, RC, t1, t0, x + q * MAP_NFEAT);
            }
            continue;
        }
        uint64_t w = a.word[i];
        RangeState st;
        st.load(s, i);
        if (WM > 0 && !isN)
            range_read_fast<RC, (WM > 0 ? WM : 1), CS>(a, i, t0, t1, last, w, st, n_probe, n_cand);
        else
            range_read_generic<RC>(a, i, t0, t1, last, w, st, n_probe, n_cand);
        a.last[i] = (uint8_t)last;
        a.word[i] = w;
        st.store(s, i);
        if (x) map_features(a, st.c, last, isN, RC, t1, t0, x + q * MAP_NFEAT);
    }
    if (a.work) {
        wave_add_u64(&a.work[1], n_probe);
        wave_add_u64(&a.work[2], n_cand);
    }
}

// Learned mode: a read whose logit is below its threshold ends its pass.
// A read that already has a mapping (best count <= max_mm) only gives up a
// lower count by stopping; one without gives up being mapped at all, so it
// has a threshold of its own.
__global__ void k_map_stop(MapState s, const uint32_t* __restrict__ act, const uint32_t* __restrict__ d_n,
                           const float* __restrict__ logit, const uint8_t* __restrict__ last, uint32_t max_mm,
                           float thr_mapped, float thr_unmapped) {
    GRID_STRIDE(q, *d_n) {
        const uint32_t i = act[q];
        if (logit[q] < (last[i] <= max_mm ? thr_mapped : thr_unmapped)) s.stop[i] = 1;
    }
}

// Reads that stay in the pass: above the target and not stopped.
struct PredMapActive {
    const uint32_t* act;
    const uint8_t* last;
    const uint8_t* stop;
    __device__ bool operator()(uint32_t q) const {
        const uint32_t i = act[q];
        return last[i] > MAP_TARGET && !stop[i];
    }
};

// Trace mode: records of sampled reads at a checkpoint (features, the best
// count there, the read's slot).
__global__ void k_map_trace_emit(const uint32_t* __restrict__ act, const uint32_t* __restrict__ d_n,
                                 const float* __restrict__ x, const uint8_t* __restrict__ last, uint32_t every,
                                 float* __restrict__ tx, uint8_t* __restrict__ tlast, uint32_t* __restrict__ tslot,
                                 unsigned long long* __restrict__ tcount, uint64_t cap) {
    GRID_STRIDE(q, *d_n) {
        const uint32_t i = act[q];
        if (i % every != 0 || last[i] <= MAP_TARGET) continue;
        const unsigned long long r = atomicAdd(tcount, 1ull);
        if (r >= cap) continue;
        for (uint32_t k = 0; k < MAP_NFEAT; ++k) tx[r * MAP_NFEAT + k] = x[(uint64_t)q * MAP_NFEAT + k];
        tlast[r] = last[i];
        tslot[r] = i;
    }
}

// Labels of the records [r0, r1) once their pass is over: 1 when the pass
// ended with a lower count than at the checkpoint.
__global__ void k_map_trace_label(const uint8_t* __restrict__ tlast, const uint32_t* __restrict__ tslot,
                                  const uint8_t* __restrict__ last, uint64_t r0, uint64_t r1, float* __restrict__ ty) {
    GRID_STRIDE(k, r1 - r0) {
        const uint64_t r = r0 + k;
        ty[r] = last[tslot[r]] < tlast[r] ? 1.f : 0.f;
    }
}

struct EmitIota {
    uint32_t* out;
    __device__ void operator()(uint32_t i, uint32_t j) const { out[j] = i; }
};

struct PredNeedsMap {
    const uint8_t* last;
    __device__ bool operator()(uint32_t i) const { return last[i] > MAP_TARGET; }
};

}  // namespace curc
