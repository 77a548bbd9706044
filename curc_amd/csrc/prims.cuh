// Device-wide primitives written for wave64 / gfx950: exclusive scan,
// stable LSD radix sort of (key, u32 value) pairs, and stable compaction
// whose element count lives in device memory (so loops over offsets need no
// host round trip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "common.cuh"

namespace curc {

#define CURC_HIP(x)                                                                          \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess)                                                                \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) +     \
                                     " at " + __FILE__ + ":" + std::to_string(__LINE__));    \
    } while (0)

// Growable device scratch owned by a context: get(slot, bytes) returns a
// buffer of at least `bytes`, reallocating only when it must grow.
// `gen` counts reallocations: work captured into a graph stays valid while
// it is unchanged.
struct Scratch {
    static constexpr int SLOTS = 256;
    void* ptr[SLOTS] = {};
    size_t cap[SLOTS] = {};
    size_t total = 0;
    uint64_t gen = 0;
    void* get(int slot, size_t bytes) {
        if (bytes == 0) bytes = 16;
        if (cap[slot] < bytes) {
            ++gen;
            if (ptr[slot]) CURC_HIP(hipFree(ptr[slot]));
            total -= cap[slot];
            size_t c = bytes + bytes / 8 + 256;
            CURC_HIP(hipMalloc(&ptr[slot], c));
            cap[slot] = c;
            total += c;
        }
        return ptr[slot];
    }
    template <typename T>
    T* as(int slot, size_t count) {
        return static_cast<T*>(get(slot, count * sizeof(T)));
    }
    void release() {
        ++gen;
        for (int i = 0; i < SLOTS; ++i) {
            if (ptr[i]) hipFree(ptr[i]);
            ptr[i] = nullptr;
            cap[i] = 0;
        }
        total = 0;
    }
};

__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = __lane_id();
    return (l == 0) ? 0ull : ((~0ull) >> (64 - l));
}

// Inclusive wave64 scan of a u32.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    uint32_t lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}

// Exclusive block scan for 256 threads (4 waves); returns the exclusive
// prefix and writes the block total to *total.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t* lds4, uint32_t* total) {
    uint32_t t = threadIdx.x, w = t >> 6, lane = t & 63;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) lds4[w] = inc;
    __syncthreads();
    uint32_t wo = 0, tot = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        uint32_t s = lds4[i];
        if (i < w) wo += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wo + inc - v;
}

// ---------------------------------------------------------------- scan ----
constexpr uint32_t SCAN_ITEMS = 8;
constexpr uint32_t SCAN_TILE = 256 * SCAN_ITEMS;

// Exclusive block scan of a T for 256 threads (4 waves).
template <typename T>
__device__ __forceinline__ T block_excl_scan256_t(T v, T* lds4, T* total) {
    uint32_t t = threadIdx.x, w = t >> 6, lane = t & 63;
    T inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T u = __shfl_up(inc, d, 64);
        if (lane >= (uint32_t)d) inc += u;
    }
    if (lane == 63) lds4[w] = inc;
    __syncthreads();
    T wo = 0, tot = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        T s = lds4[i];
        if (i < w) wo += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wo + inc - v;
}

template <typename T>
__global__ void __launch_bounds__(256) k_scan_reduce(const T* __restrict__ in, uint64_t n, T* __restrict__ sums) {
    __shared__ T lds[4];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    T s = 0;
#pragma unroll
    for (uint32_t r = 0; r < SCAN_ITEMS; ++r) {
        uint64_t i = base + r * 256 + threadIdx.x;
        if (i < n) s += in[i];
    }
    T tot;
    block_excl_scan256_t<T>(s, lds, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Exclusive scan of one tile per block plus the block offset.  Thread t owns
// elements [t*ITEMS, t*ITEMS+ITEMS) of the tile, staged through LDS so global
// loads stay coalesced.
template <typename T>
__global__ void __launch_bounds__(256) k_scan_apply(const T* __restrict__ in, T* __restrict__ out, uint64_t n,
                                                    const T* __restrict__ offs) {
    __shared__ T tile[SCAN_TILE];
    __shared__ T lds[4];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
#pragma unroll
    for (uint32_t r = 0; r < SCAN_ITEMS; ++r) {
        uint64_t i = base + r * 256 + threadIdx.x;
        tile[r * 256 + threadIdx.x] = (i < n) ? in[i] : (T)0;
    }
    __syncthreads();
    T v[SCAN_ITEMS];
    T s = 0;
#pragma unroll
    for (uint32_t r = 0; r < SCAN_ITEMS; ++r) {
        v[r] = tile[threadIdx.x * SCAN_ITEMS + r];
        s += v[r];
    }
    T tot;
    T ex = block_excl_scan256_t<T>(s, lds, &tot) + (offs ? offs[blockIdx.x] : (T)0);
#pragma unroll
    for (uint32_t r = 0; r < SCAN_ITEMS; ++r) {
        tile[threadIdx.x * SCAN_ITEMS + r] = ex;
        ex += v[r];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < SCAN_ITEMS; ++r) {
        uint64_t i = base + r * 256 + threadIdx.x;
        if (i < n) out[i] = tile[r * 256 + threadIdx.x];
    }
}

// Exclusive scan of n elements (in may equal out).  Uses scratch slots
// [slot, slot+depth).
template <typename T>
inline void scan_excl_t(const T* in, T* out, uint64_t n, Scratch& S, int slot, hipStream_t st) {
    if (n == 0) return;
    uint64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles == 1) {
        hipLaunchKernelGGL(k_scan_apply<T>, dim3(1), dim3(256), 0, st, in, out, n, (const T*)nullptr);
        return;
    }
    T* sums = S.as<T>(slot, tiles);
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3((uint32_t)tiles), dim3(256), 0, st, in, n, sums);
    scan_excl_t<T>(sums, sums, tiles, S, slot + 1, st);
    hipLaunchKernelGGL(k_scan_apply<T>, dim3((uint32_t)tiles), dim3(256), 0, st, in, out, n, (const T*)sums);
}

inline void scan_excl(const uint32_t* in, uint32_t* out, uint64_t n, Scratch& S, int slot, hipStream_t st) {
    scan_excl_t<uint32_t>(in, out, n, S, slot, st);
}

// ----------------------------------------------------------- radix sort ----
// A tile is 256 threads x ITEMS keys; the tile size is a launch shape
// (results do not depend on it), chosen per thread by radix_items().
inline uint32_t& radix_items() {
    static thread_local uint32_t items = 16;
    return items;
}

template <typename K, uint32_t ITEMS>
__global__ void __launch_bounds__(256) k_radix_hist(const K* __restrict__ keys, uint64_t n, uint32_t shift,
                                                    uint32_t* __restrict__ hist, uint32_t tiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * (256u * ITEMS);
#pragma unroll 4
    for (uint32_t r = 0; r < ITEMS; ++r) {
        uint64_t i = base + r * 256 + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter.  Wave w of a tile owns its contiguous quarter (64 x ITEMS
// keys, 64 per round) and ranks them against wave-private digit counts in
// LDS (8 ballots per round find each lane's peers), without block barriers
// between rounds; one prefix over the four waves' counts then orders the
// waves, so within a digit the order is (wave, round, lane), the input order.
// The tile is sorted by digit in LDS (each key at its digit's tile-local
// start plus its rank) and written out in that order, so consecutive threads
// write consecutive addresses of a digit's run.
template <typename K, uint32_t ITEMS>
__global__ void __launch_bounds__(256) k_radix_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                       K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                       uint64_t n, uint32_t shift,
                                                       const uint32_t* __restrict__ offs, uint32_t tiles) {
    constexpr uint32_t TILE = 256u * ITEMS;
    __shared__ K sk[TILE];
    __shared__ uint32_t sv[TILE];
    __shared__ uint32_t run[256];    // global start of this tile's run of digit d
    __shared__ uint32_t lbase[256];  // tile-local start of digit d
    __shared__ uint32_t wc[4][256];  // per-wave digit counts, then each wave's start within the digit
    __shared__ uint32_t scan4[4];
    const uint32_t t = threadIdx.x, w = t >> 6, lane = t & 63u;
    wc[0][t] = wc[1][t] = wc[2][t] = wc[3][t] = 0;
    {
        // offs is the exclusive scan of the (digit, tile) counts in digit-major
        // order, so this tile's count of digit t is the step to the next entry
        const uint64_t idx = (uint64_t)t * tiles + blockIdx.x;
        const uint32_t o = offs[idx];
        const uint32_t o1 = idx + 1 < 256ull * tiles ? offs[idx + 1] : (uint32_t)n;
        run[t] = o;
        uint32_t tot;
        lbase[t] = block_excl_scan256(o1 - o, scan4, &tot);  // its barriers also publish wc's zeros
    }
    const uint64_t base = (uint64_t)blockIdx.x * TILE + (uint64_t)w * (64u * ITEMS);
    const uint64_t lt = lanemask_lt();
    K key[ITEMS];
    uint32_t val[ITEMS], rank[ITEMS];
#pragma unroll
    for (uint32_t r = 0; r < ITEMS; ++r) {
        const uint64_t i = base + r * 64 + lane;
        const bool valid = i < n;
        key[r] = valid ? kin[i] : (K)0;
        val[r] = valid ? vin[i] : 0u;
    }
#pragma unroll
    for (uint32_t r = 0; r < ITEMS; ++r) {
        const bool valid = base + r * 64 + lane < n;
        const uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t below = (uint32_t)__popcll(peers & lt);
        const uint32_t prior = valid ? wc[w][d] : 0u;
        rank[r] = prior + below;
        // every lane of the wave has read its count before the first lane of
        // each digit group adds the group (LDS operations of a wave are
        // performed in order)
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0) wc[w][d] = prior + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    {
        uint32_t acc = 0;
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            const uint32_t c = wc[q][t];
            wc[q][t] = acc;
            acc += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < ITEMS; ++r) {
        if (base + r * 64 + lane < n) {
            const uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
            const uint32_t pos = lbase[d] + wc[w][d] + rank[r];
            sk[pos] = key[r];
            sv[pos] = val[r];
        }
    }
    __syncthreads();
    const uint64_t tbase = (uint64_t)blockIdx.x * TILE;
    const uint32_t tile_n = n - tbase < TILE ? (uint32_t)(n - tbase) : TILE;
    for (uint32_t j = t; j < tile_n; j += 256) {
        const K k = sk[j];
        const uint32_t d = (uint32_t)(k >> shift) & 255u;
        const uint32_t pos = run[d] + (j - lbase[d]);
        kout[pos] = k;
        vout[pos] = sv[j];
    }
}

// Sorts n (key, value) pairs by key bits [begin_bit, end_bit) in 8-bit
// digits, stably.  Ping-pongs between (k, v) and (kt, vt); on return k and v
// point at the sorted data.
template <typename K, uint32_t ITEMS>
void radix_sort_pairs_t(K*& k, uint32_t*& v, K*& kt, uint32_t*& vt, uint64_t n, uint32_t begin_bit,
                        uint32_t end_bit, Scratch& S, int slot, hipStream_t st) {
    constexpr uint32_t TILE = 256u * ITEMS;
    uint32_t tiles = (uint32_t)((n + TILE - 1) / TILE);
    uint32_t* hist = S.as<uint32_t>(slot, (size_t)256 * tiles);
    for (uint32_t sh = begin_bit; sh < end_bit; sh += 8) {
        hipLaunchKernelGGL((k_radix_hist<K, ITEMS>), dim3(tiles), dim3(256), 0, st, (const K*)k, n, sh, hist, tiles);
        scan_excl(hist, hist, (uint64_t)256 * tiles, S, slot + 1, st);
        hipLaunchKernelGGL((k_radix_scatter<K, ITEMS>), dim3(tiles), dim3(256), 0, st, (const K*)k,
                           (const uint32_t*)v, kt, vt, n, sh, (const uint32_t*)hist, tiles);
        K* tk = k; k = kt; kt = tk;
        uint32_t* tv = v; v = vt; vt = tv;
    }
}

template <typename K>
void radix_sort_pairs(K*& k, uint32_t*& v, K*& kt, uint32_t*& vt, uint64_t n, uint32_t begin_bit,
                      uint32_t end_bit, Scratch& S, int slot, hipStream_t st) {
    if (n <= 1) return;
    switch (radix_items()) {
        case 8: radix_sort_pairs_t<K, 8>(k, v, kt, vt, n, begin_bit, end_bit, S, slot, st); break;
        case 32: radix_sort_pairs_t<K, 32>(k, v, kt, vt, n, begin_bit, end_bit, S, slot, st); break;
        default: radix_sort_pairs_t<K, 16>(k, v, kt, vt, n, begin_bit, end_bit, S, slot, st); break;
    }
}

// ------------------------------------------------------- compaction ------
// Stable compaction over [0, *d_n) with a fixed grid: block b owns a
// contiguous chunk, so output order equals input order.
constexpr uint32_t COMPACT_GRID = 1024;
// Lists whose host-known bound is at most COMPACT_SMALL go through one
// workgroup in one launch (16 items per thread per round).
constexpr uint32_t COMPACT_SMALL = 1u << 12;
constexpr uint32_t COMPACT1_ITEMS = 16;

template <typename Pred>
__global__ void __launch_bounds__(256) k_compact_count(const uint32_t* __restrict__ d_n, Pred pred,
                                                       uint32_t* __restrict__ counts) {
    __shared__ uint32_t lds[4];
    uint32_t n = *d_n;
    uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
    uint32_t lo = blockIdx.x * chunk;
    uint32_t hi = min(n, lo + chunk);
    uint32_t c = 0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += 256) c += pred(i) ? 1u : 0u;
    uint32_t tot;
    block_excl_scan256(c, lds, &tot);
    if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_compact_offsets(uint32_t* __restrict__ counts, uint32_t nblk,
                                                          uint32_t* __restrict__ d_out_n) {
    // nblk <= 1024: one element per thread.
    __shared__ uint32_t lds[16];
    uint32_t t = threadIdx.x, w = t >> 6, lane = t & 63;
    uint32_t v = (t < nblk) ? counts[t] : 0u;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    uint32_t wo = 0, tot = 0;
    for (uint32_t i = 0; i < 16; ++i) {
        if (i < w) wo += lds[i];
        tot += lds[i];
    }
    if (t < nblk) counts[t] = wo + inc - v;
    if (t == 0) *d_out_n = tot;
}

template <typename Pred, typename Emit>
__global__ void __launch_bounds__(256) k_compact_scatter(const uint32_t* __restrict__ d_n, Pred pred, Emit emit,
                                                         const uint32_t* __restrict__ offs) {
    __shared__ uint32_t lds[4];
    uint32_t n = *d_n;
    uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
    uint32_t lo = blockIdx.x * chunk;
    uint32_t hi = min(n, lo + chunk);
    uint32_t o = offs[blockIdx.x];
    for (uint32_t i0 = lo; i0 < hi; i0 += 256) {
        uint32_t i = i0 + threadIdx.x;
        bool f = (i < hi) && pred(i);
        uint32_t tot;
        uint32_t ex = block_excl_scan256(f ? 1u : 0u, lds, &tot);
        if (f) emit(i, o + ex);
        o += tot;
    }
}

// One workgroup of 1024 threads; each round covers 16 consecutive items per
// thread.  Correct for any n, fast for short lists.
template <typename Pred, typename Emit>
__device__ __forceinline__ void compact_small_body(const uint32_t* __restrict__ d_n, Pred pred, Emit emit,
                                                   uint32_t* __restrict__ d_out_n) {
    __shared__ uint32_t lds[16];
    const uint32_t n = *d_n;
    const uint32_t t = threadIdx.x, w = t >> 6, lane = t & 63;
    uint32_t o = 0;
    for (uint64_t base = 0; base < n; base += 1024ull * COMPACT1_ITEMS) {
        const uint64_t first = base + (uint64_t)t * COMPACT1_ITEMS;
        uint32_t bits = 0;
#pragma unroll
        for (uint32_t r = 0; r < COMPACT1_ITEMS; ++r) {
            const uint64_t i = first + r;
            if (i < n && pred((uint32_t)i)) bits |= 1u << r;
        }
        const uint32_t c = (uint32_t)__popc(bits);
        const uint32_t inc = wave_incl_scan(c);
        if (lane == 63) lds[w] = inc;
        __syncthreads();
        uint32_t wo = 0, tot = 0;
#pragma unroll
        for (uint32_t i = 0; i < 16; ++i) {
            const uint32_t s = lds[i];
            if (i < w) wo += s;
            tot += s;
        }
        __syncthreads();
        uint32_t j = o + wo + inc - c;
        while (bits) {
            const uint32_t r = (uint32_t)__builtin_ctz(bits);
            bits &= bits - 1;
            emit((uint32_t)(first + r), j++);
        }
        o += tot;
    }
    if (t == 0) *d_out_n = o;
}

template <typename Pred, typename Emit>
__global__ void __launch_bounds__(1024) k_compact_small(const uint32_t* __restrict__ d_n, Pred pred, Emit emit,
                                                        uint32_t* __restrict__ d_out_n) {
    compact_small_body(d_n, pred, emit, d_out_n);
}

// ------------------------------------------- two lists in one launch ----
// Stable compaction of two lists in a single launch (decoupled look-back).
// The grid holds tilesA tiles of list A then tilesB tiles of list B, each
// tile COMPACT2_TILE items; workgroup b processes tile b.  Each tile
// publishes its count (AGG) at once and its inclusive prefix (INC) after
// looking back over its predecessors.  Workgroups are dispatched in
// increasing id order on each XCD, so a wait cannot close a cycle: the
// lowest undispatched id waits only for slots held by lower ids, whose own
// waits end.  Status words hold state and value together, so one atomic
// load sees a consistent pair; nothing else is published through them, so
// the loads and stores are relaxed (no cache invalidation while a wave
// waits).  The status words must be zero at launch: the caller clears them
// between launches (on the same stream) and the launch leaves them set.
// The look-back wait is bounded; a wait that runs out sets st->err and the
// launch's output is then invalid (the caller checks).
constexpr uint32_t COMPACT2_ITEMS = 16;
constexpr uint32_t COMPACT2_TILE = 256 * COMPACT2_ITEMS;
constexpr uint32_t COMPACT2_MAX_TILES = 16384;  // per launch (both lists)
constexpr uint64_t LB_AGG = 1ull << 32, LB_INC = 2ull << 32;
constexpr uint32_t LB_SPIN_LIMIT = 1u << 26;

struct LookbackState {
    uint32_t err, pad[3];
    uint64_t status[COMPACT2_MAX_TILES];
};

template <typename Pred, typename Emit>
__device__ __forceinline__ void compact2_tile(const uint32_t* __restrict__ d_n, Pred pred, Emit emit,
                                              uint32_t* __restrict__ d_out_n, uint64_t* __restrict__ status,
                                              uint32_t tile, uint32_t* __restrict__ err, uint32_t* lds) {
    const uint32_t n = *d_n;
    const uint64_t first = (uint64_t)tile * COMPACT2_TILE;
    if (tile > 0 && first >= n) return;  // past the list: no later tile needs it
    // round r covers items first + r*256 + [0, 256): reads and writes are
    // coalesced, and an item's rank in the tile follows (round, wave, lane)
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    const uint64_t lt = lanemask_lt();
    uint64_t bal[COMPACT2_ITEMS];
#pragma unroll
    for (uint32_t r = 0; r < COMPACT2_ITEMS; ++r) {
        const uint64_t i = first + r * 256 + threadIdx.x;
        bal[r] = __ballot(i < n && pred((uint32_t)i));
        if (lane == 0) lds[8 + r * 4 + w] = (uint32_t)__popcll(bal[r]);
    }
    __syncthreads();
    uint32_t agg = 0;
    if (threadIdx.x < 64) {
        // exclusive offsets of the 64 (round, wave) groups, and the tile total
        const uint32_t v = lds[8 + threadIdx.x];
        const uint32_t inc = wave_incl_scan(v);
        lds[8 + 64 + threadIdx.x] = inc - v;
        agg = __shfl(inc, 63);
    }
    if (threadIdx.x < 64) {
        // wave 0 looks back 64 predecessors at a time; lane l reads tile
        // top - l, and the nearest inclusive prefix ends the walk
        uint32_t excl = 0;
        if (tile == 0) {
            if (lane == 0) __hip_atomic_store(&status[0], LB_INC | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0) __hip_atomic_store(&status[tile], LB_AGG | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int64_t top = (int64_t)tile - 1;
            for (;;) {
                const int64_t j = top - (int64_t)lane;
                uint64_t v = LB_INC;  // before tile 0: an inclusive prefix of 0
                if (j >= 0) {
                    uint32_t spins = 0;
                    while (((v = __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >> 32) ==
                           0) {
                        if (++spins == LB_SPIN_LIMIT) {
                            atomicExch(err, 1u);
                            v = LB_INC;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                const uint64_t inc = __ballot((v >> 32) == 2);
                const uint32_t stop = inc ? (uint32_t)(__ffsll((long long)inc) - 1) : 63u;
                const uint32_t part = wave_incl_scan(lane <= stop ? (uint32_t)v : 0u);
                excl += __shfl(part, 63);
                if (inc) break;
                top -= 64;
            }
            if (lane == 0)
                __hip_atomic_store(&status[tile], LB_INC | (uint64_t)(excl + agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) {
            lds[4] = excl;
            // the tile holding the list's last item (or tile 0 of an empty list) reports the total
            if (first + COMPACT2_TILE >= n) *d_out_n = excl + agg;
        }
    }
    __syncthreads();
    const uint32_t base = lds[4];
#pragma unroll
    for (uint32_t r = 0; r < COMPACT2_ITEMS; ++r) {
        if ((bal[r] >> lane) & 1ull)
            emit((uint32_t)(first + r * 256 + threadIdx.x),
                 base + lds[8 + 64 + r * 4 + w] + (uint32_t)__popcll(bal[r] & lt));
    }
}

template <typename PredA, typename EmitA, typename PredB, typename EmitB>
__global__ void __launch_bounds__(256) k_compact2(const uint32_t* __restrict__ na, PredA pa, EmitA ea,
                                                  uint32_t* __restrict__ outa, uint32_t tiles_a,
                                                  const uint32_t* __restrict__ nb, PredB pb, EmitB eb,
                                                  uint32_t* __restrict__ outb, LookbackState* __restrict__ st) {
    __shared__ uint32_t lds[8 + 128];
    const uint32_t t = blockIdx.x;
    if (t < tiles_a)
        compact2_tile(na, pa, ea, outa, st->status, t, &st->err, lds);
    else
        compact2_tile(nb, pb, eb, outb, st->status + tiles_a, t - tiles_a, &st->err, lds);
}

// Compacts list A and list B (host-known bounds bound_a, bound_b) in one
// launch.  st->status[0, compact2_words(bound_a, bound_b) / 2) must be zero
// at launch; the launch leaves it set.
inline uint64_t compact2_tiles(uint64_t bound) {
    return std::max<uint64_t>(1, (bound + COMPACT2_TILE - 1) / COMPACT2_TILE);
}
// 32-bit words of status the launch sets
inline uint64_t compact2_words(uint64_t bound_a, uint64_t bound_b) {
    return 2 * (compact2_tiles(bound_a) + compact2_tiles(bound_b));
}

template <typename PredA, typename EmitA, typename PredB, typename EmitB>
void compact2(const uint32_t* na, PredA pa, EmitA ea, uint32_t* outa, uint64_t bound_a, const uint32_t* nb,
              PredB pb, EmitB eb, uint32_t* outb, uint64_t bound_b, LookbackState* st, hipStream_t s) {
    const uint32_t ta = (uint32_t)compact2_tiles(bound_a);
    const uint32_t tb = (uint32_t)compact2_tiles(bound_b);
    if ((uint64_t)ta + tb > COMPACT2_MAX_TILES) throw std::length_error("compact2: lists too long");
    hipLaunchKernelGGL((k_compact2<PredA, EmitA, PredB, EmitB>), dim3(ta + tb), dim3(256), 0, s, na, pa, ea, outa,
                       ta, nb, pb, eb, outb, st);
}

// Stable compaction: emit(i, j) is called for the j-th i in [0, *d_n) with
// pred(i); the number of emitted items is written to *d_out_n (device).
// `bound` is a host-known upper bound of *d_n.
template <typename Pred, typename Emit>
void compact(const uint32_t* d_n, Pred pred, Emit emit, uint32_t* d_out_n, Scratch& S, int slot,
             hipStream_t st, uint64_t bound = ~0ull) {
    if (bound <= COMPACT_SMALL) {
        hipLaunchKernelGGL((k_compact_small<Pred, Emit>), dim3(1), dim3(1024), 0, st, d_n, pred, emit, d_out_n);
        return;
    }
    const uint32_t grid = COMPACT_GRID;
    uint32_t* counts = S.as<uint32_t>(slot, grid);
    hipLaunchKernelGGL(k_compact_count<Pred>, dim3(grid), dim3(256), 0, st, d_n, pred, counts);
    hipLaunchKernelGGL(k_compact_offsets, dim3(1), dim3(1024), 0, st, counts, grid, d_out_n);
    hipLaunchKernelGGL((k_compact_scatter<Pred, Emit>), dim3(grid), dim3(256), 0, st, d_n, pred, emit,
                       (const uint32_t*)counts);
}

}  // namespace curc
