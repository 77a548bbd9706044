// HIP kernels of the CURC pseudogenome path (gfx950, wave64).
//
// Design notes (the reference runs the same stages as thrust calls over
// __device__ lambdas, compress.cu:1349-1722):
//  * every loop over offsets keeps its list lengths in device memory, so a
//    whole loop is enqueued without host round trips;
//  * the reference's racy first-come claims (compress.cu:1518) and the
//    merge-join of full_match_gpu (compress.cu:1089-1160) are both computed
//    here as serial dictatorship inside each group of equal overlap strings:
//    a suffix proposes its first free candidate; a group where two suffixes
//    propose the same prefix is handed to one leader thread that replays the
//    group in the reference's suffix order.
#pragma once
#include "common.cuh"
#include "prims.cuh"

namespace curc {

#define GRID_STRIDE(i, n) \
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

// Slot of this lane's entry in a list that the active lanes of the
// wavefront that set `want` append to, with one atomic per wavefront (every
// active lane calls it).
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool want) {
    const uint64_t act = __ballot(1);
    const uint64_t bal = __ballot(want);
    const int leader = __builtin_ctzll(act);
    uint32_t base = 0;
    if (bal && (int)(threadIdx.x & 63u) == leader) base = atomicAdd(counter, (uint32_t)__popcll(bal));
    base = __shfl(base, leader, 64);
    return base + (uint32_t)__popcll(bal & lanemask_lt());
}

// Plain device copy in 16-byte words (the measured bandwidth ceiling the
// benchmark reports beside the vendor peak).
__global__ void __launch_bounds__(256) k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     uint64_t n) {
    GRID_STRIDE(i, n) dst[i] = src[i];
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, uint64_t n, T v) {
    GRID_STRIDE(i, n) p[i] = v;
}

__global__ void k_iota(uint32_t* __restrict__ p, uint64_t n) {
    GRID_STRIDE(i, n) p[i] = (uint32_t)i;
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

// ------------------------------------------------------------------ pack ---
// The 32 bytes from c (nb <= 32 of them wanted) as eight little-endian
// dwords, read as aligned dwords; a dword reaching outside [begin, end) is
// read byte by byte.  Bytes past nb are left as loaded (or zero).
__device__ __forceinline__ void load_bytes32(const uint8_t* c, uint32_t nb, const uint8_t* begin, const uint8_t* end,
                                             uint32_t (&a)[8]) {
    const uintptr_t p = (uintptr_t)c;
    const uint32_t sh = (uint32_t)(p & 3u);
    const uint8_t* base = c - sh;
    uint32_t d[9];
#pragma unroll
    for (uint32_t k = 0; k < 9; ++k) {
        const uint8_t* q = base + 4 * k;
        d[k] = 0u;
        if (4 * k < sh + nb) {
            if (q >= begin && q + 4 <= end) {
                d[k] = *reinterpret_cast<const uint32_t*>(q);
            } else {
#pragma unroll
                for (uint32_t b = 0; b < 4; ++b)
                    if (q + b >= begin && q + b < end) d[k] |= (uint32_t)q[b] << (8 * b);
            }
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) a[k] = __builtin_amdgcn_alignbyte(d[k + 1], d[k], sh);
}

// ASCII reads to 2-bit words, one thread per packed word (enc_base & 3 per
// base, first base in the top bits).  A read holding an uppercase 'N' gets
// isN set (isN cleared beforehand; the flag byte is set with an atomic OR
// on its dword, which few reads need).
__global__ void __launch_bounds__(256) k_pack_w(const uint8_t* __restrict__ ascii, uint64_t n, uint32_t L, uint32_t W,
                                                uint64_t* __restrict__ packed, uint8_t* __restrict__ isN) {
    const uint8_t* end = ascii + n * L;
    GRID_STRIDE(t, n * W) {
        const uint64_t r = t / W;
        const uint32_t w = (uint32_t)(t - r * W);
        const uint32_t b0 = 32u * w;
        const uint32_t nb = min(32u, L - b0);
        uint32_t a[8];
        load_bytes32(ascii + r * L + b0, nb, ascii, end, a);
        uint64_t v = 0;
        uint32_t hasN = 0;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) {
            // 'N' among this dword's wanted bytes (zero-byte test; others forced non-zero)
            const uint32_t valid = 4 * k < nb ? min(4u, nb - 4 * k) : 0u;
            const uint32_t keep = valid >= 4 ? 0xFFFFFFFFu : ((1u << (8 * valid)) - 1u);
            const uint32_t x = (a[k] ^ 0x4E4E4E4Eu) | ~keep;
            hasN |= (x - 0x01010101u) & ~x & 0x80808080u;
#pragma unroll
            for (uint32_t b = 0; b < 4; ++b) {
                const uint32_t j = 4 * k + b;
                if (j < nb) {
                    const uint32_t u = (a[k] >> (8 * b)) & 0xDFu;
                    const uint32_t code = (u == 'C' ? 1u : 0u) | (u == 'G' ? 2u : 0u) | (u == 'T' ? 3u : 0u);
                    v |= (uint64_t)code << (62u - 2u * j);
                }
            }
        }
        packed[t] = v;
        if (isN && hasN) {
            const uintptr_t fp = (uintptr_t)(isN + r);
            atomicOr(reinterpret_cast<uint32_t*>(fp & ~(uintptr_t)3), 1u << (8 * (fp & 3)));
        }
    }
}

struct PredIsN {
    const uint8_t* isN;
    bool want;
    __device__ bool operator()(uint32_t i) const { return (isN[i] != 0) == want; }
};
struct EmitIndex {
    uint32_t* out;
    __device__ void operator()(uint32_t i, uint32_t j) const { out[j] = i; }
};

__global__ void k_gather_ascii(const uint8_t* __restrict__ ascii, const uint32_t* __restrict__ ids,
                               uint64_t n, uint32_t L, uint8_t* __restrict__ out) {
    GRID_STRIDE(b, n * L) {
        uint64_t r = b / L, j = b - r * L;
        out[b] = ascii[(uint64_t)ids[r] * L + j];
    }
}

__global__ void k_gather_packed(const uint64_t* __restrict__ R, const uint32_t* __restrict__ ids,
                                uint64_t n, uint32_t W, uint64_t* __restrict__ out) {
    GRID_STRIDE(t, n * W) {
        uint64_t r = t / W, w = t - r * W;
        out[t] = R[(uint64_t)ids[r] * W + w];
    }
}

// ----------------------------------------------------------------- dedup ---
__global__ void k_gather_word(const uint64_t* __restrict__ R, const uint32_t* __restrict__ ids, uint64_t n,
                              uint32_t W, uint32_t w, uint64_t* __restrict__ keys) {
    GRID_STRIDE(i, n) keys[i] = R[(uint64_t)ids[i] * W + w];
}

// Tie refinement of the dedup sort.  After a stable sort of all reads by
// their first word, only reads in runs of equal first words need the other
// words: their positions are compacted, the subset is sorted by the
// remaining words (least significant first) and then by run number, all
// stably, and written back into the same positions.  The order equals a
// full lexicographic sort with ties by ascending id.
struct PredTied {
    const uint64_t* key;  // first words in sorted order
    uint64_t n;
    uint64_t mask;        // the bits the sort ordered by
    __device__ bool operator()(uint32_t i) const {
        const uint64_t k = key[i] & mask;
        return (i > 0 && (key[i - 1] & mask) == k) || (i + 1 < n && (key[i + 1] & mask) == k);
    }
};

// Subset entry j: its read id, its index (the value the subset sorts
// carry) and 1 when it starts a run.
__global__ void k_tie_setup(const uint32_t* __restrict__ pos, const uint32_t* __restrict__ d_m,
                            const uint64_t* __restrict__ key, uint64_t mask, const uint32_t* __restrict__ ids,
                            uint32_t* __restrict__ sid, uint32_t* __restrict__ perm, uint32_t* __restrict__ head) {
    const uint32_t m = *d_m;
    GRID_STRIDE(j, m) {
        const uint32_t p = pos[j];
        sid[j] = ids[p];
        perm[j] = (uint32_t)j;
        head[j] = (j == 0 || (key[pos[j - 1]] & mask) != (key[p] & mask)) ? 1u : 0u;
    }
}

// Run number of entry j from the exclusive scan of the run-start flags.
__global__ void k_tie_runid(const uint32_t* __restrict__ d_m, const uint32_t* __restrict__ head,
                            const uint32_t* __restrict__ excl, uint32_t* __restrict__ run) {
    const uint32_t m = *d_m;
    GRID_STRIDE(j, m) run[j] = excl[j] + head[j] - 1u;
}

// Word w of the read behind each permuted subset entry.
__global__ void k_gather_word_perm(const uint64_t* __restrict__ R, const uint32_t* __restrict__ sid,
                                   const uint32_t* __restrict__ perm, uint64_t m, uint32_t W, uint32_t w,
                                   uint64_t* __restrict__ keys) {
    GRID_STRIDE(q, m) keys[q] = R[(uint64_t)sid[perm[q]] * W + w];
}

__global__ void k_gather_u32_perm(const uint32_t* __restrict__ src, const uint32_t* __restrict__ perm, uint64_t m,
                                  uint32_t* __restrict__ keys) {
    GRID_STRIDE(q, m) keys[q] = src[perm[q]];
}

__global__ void k_tie_scatter(const uint32_t* __restrict__ pos, const uint32_t* __restrict__ sid,
                              const uint32_t* __restrict__ perm, uint64_t m, uint32_t* __restrict__ ids) {
    GRID_STRIDE(q, m) ids[pos[q]] = sid[perm[q]];
}

// compress.cu:1409-1420: adjacent equal reads in sorted order are linked.
// eqf[i]: the read at dedup position i equals the one at i + 1 (0 at n - 1),
// so that the reads without a predecessor or successor in dedup order are
// found from a sequential pass (PredNoPred / PredNoSucc) instead of gathers
// of next and prev.
__global__ void k_dedup_link(const uint64_t* __restrict__ R, const uint32_t* __restrict__ sorted, uint64_t n,
                             uint32_t W, uint32_t* __restrict__ next, uint32_t* __restrict__ prev,
                             uint16_t* __restrict__ offset, uint8_t* __restrict__ eqf) {
    if (blockIdx.x == 0 && threadIdx.x == 0) eqf[n - 1] = 0;
    GRID_STRIDE(i, n - 1) {
        uint32_t a = sorted[i], b = sorted[i + 1];
        const uint64_t* ra = R + (uint64_t)a * W;
        const uint64_t* rb = R + (uint64_t)b * W;
        bool eq = true;
        for (uint32_t w = 0; w < W; ++w)
            if (ra[w] != rb[w]) {
                eq = false;
                break;
            }
        eqf[i] = eq ? 1 : 0;
        if (eq) {
            next[a] = b;
            prev[b] = a;
            offset[a] = 0;
        }
    }
}

// Dedup positions without a predecessor / successor (the dedup links are the
// only links when the prefix and suffix lists are formed).
struct PredNoPred {
    const uint8_t* eqf;
    __device__ bool operator()(uint32_t i) const { return i == 0 || eqf[i - 1] == 0; }
};
struct PredNoSucc {
    const uint8_t* eqf;
    __device__ bool operator()(uint32_t i) const { return eqf[i] == 0; }
};

struct PredLinkState {
    const uint32_t* list;
    const uint32_t* link;  // next or prev
    __device__ bool operator()(uint32_t i) const { return link[list[i]] == INV; }
};
struct EmitGather {
    const uint32_t* list;
    uint32_t* out;
    __device__ void operator()(uint32_t i, uint32_t j) const { out[j] = list[i]; }
};
struct EmitGatherKey {
    const uint32_t* list;
    const uint64_t* R;
    uint32_t W;
    uint32_t* out;
    uint64_t* key;
    __device__ void operator()(uint32_t i, uint32_t j) const {
        uint32_t id = list[i];
        out[j] = id;
        key[j] = R[(uint64_t)id * W];
    }
};

// -------------------------------------------------------- prefix index ---
// Bucket of each prefix's k-mer; with filt, also the k-mer filter bit
// (hash bits 32..63 under fmask): a suffix whose k-mer bit is clear has no
// prefix with that k-mer, hence none equal to it, at any offset.
// Minimizer-blocked prefix k-mer filter.  A k-mer's two filter bits
// (ov_fbits) lie in the 512-bit block chosen by its
// minimizer, the least ov_mhash over its OVM-base substrings.  Neighbouring
// k-mers of a read usually share their minimizer, so the k-mers of one
// suffix at all offsets touch a few 64-byte blocks.  Exact like the plain
// filter: a clear bit proves that no prefix has the k-mer.
constexpr uint32_t OVM = 16;
constexpr uint32_t OVB_WORDS = 16;  // 512-bit blocks
constexpr uint32_t OVB_MASK = OVB_WORDS * 32 - 1;
// order of the m-mers for minimizers: a bijection of 32-bit values
__device__ __forceinline__ uint32_t ov_mhash(uint32_t v) { return v * 0x9E3779B1u; }
// the k-mer's two bit positions in its block: from bits 0..9 and 10..19
__device__ __forceinline__ uint32_t ov_fbits(uint64_t K) { return (uint32_t)((K * 0xC2B2AE3D27D4EB4Full) >> 44); }
__device__ __forceinline__ uint32_t ov_block(uint32_t mn, uint32_t bmask) {
    return (uint32_t)(((uint64_t)(mn ^ 0x5bd1e995u) * 0xD6E8FEB86659FD93ull) >> 40) & bmask;
}
// minimizer of the k-base k-mer K (2k bits, first base highest)
__device__ __forceinline__ uint32_t ov_minimizer(uint64_t K, uint32_t k) {
    uint32_t mn = 0xFFFFFFFFu;
    for (uint32_t t = 0; t + OVM <= k; ++t) mn = min(mn, ov_mhash((uint32_t)(K >> (2 * (k - OVM - t)))));
    return mn;
}

// Bucket hash of each prefix's k-mer at kip, and its bits in the plain
// filter (filt) or the minimizer-blocked one (mfilt, bmask + 1 blocks).
__global__ void k_prefix_hash(const uint64_t* __restrict__ R, const uint32_t* __restrict__ ids, uint64_t n,
                              uint32_t W, uint32_t kip, uint32_t k, uint32_t mask, uint32_t* __restrict__ hk,
                              uint32_t* __restrict__ filt, uint32_t fmask, uint32_t* __restrict__ mfilt,
                              uint32_t bmask) {
    GRID_STRIDE(i, n) {
        const uint64_t K = xw(R + (uint64_t)ids[i] * W, kip, k);
        const uint64_t m = murmur64(K);
        hk[i] = (uint32_t)(m & mask);
        if (filt) {
            const uint32_t f = (uint32_t)(m >> 32) & fmask;
            atomicOr(&filt[f >> 5], 1u << (f & 31u));
        }
        if (mfilt) {
            uint32_t* fb = mfilt + (uint64_t)ov_block(ov_minimizer(K, k), bmask) * OVB_WORDS;
            const uint32_t f = ov_fbits(K), b1 = f & OVB_MASK, b2 = (f >> 10) & OVB_MASK;
            atomicOr(&fb[b1 >> 5], 1u << (b1 & 31u));
            atomicOr(&fb[b2 >> 5], 1u << (b2 & 31u));
        }
    }
}

// CSR offsets from sorted keys: kr[b] = first index with key >= b, for
// b in [0, nb]; every entry is written by exactly one element.
__global__ void k_gap_fill(const uint32_t* __restrict__ keys, uint64_t n, uint64_t nb, uint32_t* __restrict__ kr) {
    GRID_STRIDE(i, n + 1) {
        uint64_t lo = (i == 0) ? 0 : (uint64_t)keys[i - 1] + 1;
        uint64_t hi = (i == n) ? nb : (uint64_t)keys[i];
        for (uint64_t b = lo; b <= hi; ++b) kr[b] = (uint32_t)i;
    }
}

__global__ void k_gather_key0(const uint64_t* __restrict__ R, const uint32_t* __restrict__ ids, uint64_t n,
                              uint32_t W, uint64_t* __restrict__ key) {
    GRID_STRIDE(i, n) key[i] = R[(uint64_t)ids[i] * W];
}

__global__ void k_gather_key0_dev(const uint64_t* __restrict__ R, const uint32_t* __restrict__ ids,
                                  const uint32_t* __restrict__ d_n, uint32_t W, uint64_t* __restrict__ key) {
    const uint32_t n = *d_n;
    GRID_STRIDE(i, n) key[i] = R[(uint64_t)ids[i] * W];
}

// Base encodings the exact-matching kernels run over.  Enc2: ACGT reads,
// 32 bases per word, 2 bits MSB-first (the packed reads).  Enc3: N-reads over
// the alphabet A<C<G<N<T (compress.cu:1205), 21 bases per word, 3 bits
// MSB-first below an unused top bit, so numeric order is ASCII order.
struct Enc2 {
    static constexpr uint32_t BPW = 32;
    __device__ static uint64_t word(const uint64_t* rd, uint32_t off, uint32_t n) { return xw(rd, off, n); }
    // first n bases of a read from its first word
    __device__ static uint64_t head(uint64_t w0, uint32_t n) { return w0 >> ((32u - n) << 1); }
    // n right-aligned bases in reverse order
    __device__ static uint64_t reversed(uint64_t x, uint32_t n) {
        uint64_t mask = (n == 32) ? ~0ull : ((1ull << (2 * n)) - 1);
        return revcomp(x, n) ^ mask;
    }
};

struct Enc3 {
    static constexpr uint32_t BPW = 21;
    __device__ static uint64_t word(const uint64_t* rd, uint32_t off, uint32_t n) {
        uint32_t o1 = off / 21, o2 = off - o1 * 21;
        uint64_t w = ld64(rd + o1);
        if (o2 + n <= 21) return (w >> (3 * (21 - o2 - n))) & ((1ull << (3 * n)) - 1);
        uint32_t k1 = 21 - o2, k2 = n - k1;
        uint64_t hi = w & ((1ull << (3 * k1)) - 1);
        return (hi << (3 * k2)) | (ld64(rd + o1 + 1) >> (3 * (21 - k2)));
    }
    __device__ static uint64_t head(uint64_t w0, uint32_t n) { return w0 >> (3 * (21 - n)); }
    __device__ static uint64_t reversed(uint64_t x, uint32_t n) {
        uint64_t r = 0;
        for (uint32_t t = 0; t < n; ++t) r = (r << 3) | ((x >> (3 * t)) & 7u);
        return r;
    }
};

template <uint32_t WM>
__device__ __forceinline__ uint64_t pick(const uint64_t (&a)[WM], uint32_t i) {
    uint64_t v = 0;
#pragma unroll
    for (uint32_t k = 0; k < WM; ++k) v = (i == k) ? a[k] : v;
    return v;
}

// n (1..32) bases from base `off` of a read held in registers (as xw).
template <uint32_t WM>
__device__ __forceinline__ uint64_t xw_pick(const uint64_t (&w)[WM], uint32_t off, uint32_t n) {
    const uint32_t o1 = off >> 5;
    const uint32_t o2 = (off & 31u) << 1;
    const uint64_t left = pick(w, o1) << o2;
    if ((64u - o2) >= (n << 1)) return left >> ((32u - n) << 1);
    const uint64_t right = (pick(w, o1 + 1) >> (62u - o2)) >> 2;
    return (left | right) >> ((32u - n) << 1);
}

// Overlap filter verdicts for every offset in one pass (minimizer-blocked
// filter): bit i of vb[(off - 1) * nw + i / 64] is set when the k-mer of
// suffix i at kip + off has both of its bits set, for off in [1, max_off).
// Also the exact-matching verdicts (kip 0, k 32: the suffix's 32 bases from
// off against the prefixes' first 32 bases).
// Thread i holds its read's words (its row of the packed reads R, through
// the suffix list sl), and every lane does the same work per offset, so
// lanes do not diverge.  For k = 32
// (K32) offsets go in chunks of NM = 17, the m-mers per k-mer: with the 64
// bases from the chunk's first k-mer in two words, each k-mer and m-mer of
// the chunk is a fixed shift of them, and the least m-mer hash of each
// window is min(suffix minimum over the chunk's m-mers from its start,
// prefix minimum over the next chunk's m-mers up to its end).  For other k
// the minimizer is recomputed from each k-mer.  The filter block of the
// current minimizer is kept in the thread's own LDS slot and reloaded only
// when the minimizer changes: between two offsets of one thread the other
// resident threads touch too many blocks for the caches to keep it.  Each
// 64-suffix word is written by its wave with one ballot per offset.

// p[0:len) vs s[off:off+len), with p's first word given and s's first
// min(BPW,len) bases pre-extracted in sx0.
template <class E>
__device__ __forceinline__ int cmp_ps(const uint64_t* __restrict__ R, uint32_t W, uint32_t p, uint64_t pkey,
                                      const uint64_t* __restrict__ rs, uint64_t sx0, uint32_t off, uint32_t len) {
    uint32_t n = min(E::BPW, len);
    uint64_t a = E::head(pkey, n);
    if (a != sx0) return a < sx0 ? -1 : 1;
    const uint64_t* rp = R + (uint64_t)p * W;
    for (uint32_t j = E::BPW; j < len; j += E::BPW) {
        n = min(E::BPW, len - j);
        uint64_t x = E::word(rp, j, n), y = E::word(rs, off + j, n);
        if (x != y) return x < y ? -1 : 1;
    }
    return 0;
}

// Order of two suffixes of one full-match group (equal s[off:L]): the
// reference's suffix list is sorted by read[off:] then by base off-1,
// off-2, ..., 0 (compress.cu:1013-1059; full_match_cpu keeps the same order
// with its stable multiway merge, compress.cu:1230-1242), i.e. by the
// reversed s[0:off).
template <class E>
__device__ __forceinline__ int cmp_rev_prefix(const uint64_t* __restrict__ R, uint32_t W, uint32_t a,
                                              uint32_t b, uint32_t off) {
    const uint64_t* ra = R + (uint64_t)a * W;
    const uint64_t* rb = R + (uint64_t)b * W;
    for (uint32_t t = 0; t < off; t += E::BPW) {
        uint32_t n = min(E::BPW, off - t);
        uint32_t st = off - t - n;
        uint64_t x = E::reversed(E::word(ra, st, n), n);
        uint64_t y = E::reversed(E::word(rb, st, n), n);
        if (x != y) return x < y ? -1 : 1;
    }
    return 0;
}

// State shared by the overlap (a6) and full-match (a8) claim rounds.
struct ClaimState {
    const uint64_t* R;
    uint32_t W, L;
    const uint32_t* sl;     // suffix list (ids)
    const uint32_t* d_ns;   // suffix count (device)
    const uint32_t* pl;     // prefix array (ids): a6 hash-ordered, a8 lexicographic list
    const uint64_t* pk;     // first word of each prefix
    const uint32_t* d_np;   // prefix count (device)
    uint32_t* next;
    uint32_t* prev;
    uint16_t* offset;
    uint32_t* choice;       // per suffix index: chosen prefix position or INV
    uint32_t* grp;          // per suffix index: group id (first equal position)
    uint32_t* owner;        // per prefix position: proposing suffix index
    uint32_t* contended;    // per prefix position (group id): 1 if contended
    uint32_t* head;         // per group id: member list head
    uint32_t* memb;         // per suffix index: next member
    uint32_t* res_list;
    uint32_t* res_n;
    uint32_t* pool;         // member lists of contended groups, carved per leader
    uint32_t* pool_n;
    uint64_t* kpool;        // a8: first reversed-prefix word of each pooled member (or null)
    uint8_t* slinked;       // per suffix index: linked at this offset
    uint8_t* plinked;       // per prefix position: linked at this offset (a8 only)
    uint64_t* work;         // a8: suffix-list lengths summed over offsets (or null)
    uint32_t* nf;           // a8 over lists kept in place: prefix skip pointers (see fm_find), or null
    uint64_t* slbits;       // a6 probe by verdict words: bit si set once suffix si is linked, or null
    __device__ __forceinline__ void link(uint32_t si, uint32_t px, uint32_t off) const {
        const uint32_t s = sl[si], p = pl[px];
        next[s] = p;
        prev[p] = s;
        offset[s] = (uint16_t)off;
        slinked[si] = 1;
        if (plinked) plinked[px] = 1;
        if (nf) nf[px] = px + 1;
        if (slbits) atomicOr((unsigned long long*)&slbits[si >> 6], 1ull << (si & 63u));
    }
};

struct PredFlagClear {
    const uint8_t* flags;
    __device__ bool operator()(uint32_t i) const { return flags[i] == 0; }
};

// Prefix-index bucket slots: one 64-byte line per hash bucket holding the
// bucket's size and start in the hash-ordered prefix arrays, and the ids
// and first words of its first PSLOT_N entries, so that a probe of a small
// bucket reads one line instead of the CSR offsets plus the arrays.
constexpr uint32_t PSLOT_N = 4;
struct alignas(64) PrefixSlot {
    uint32_t cnt, lo;
    uint32_t id[PSLOT_N];
    uint64_t key[PSLOT_N];
    uint32_t pad[2];
};
static_assert(sizeof(PrefixSlot) == 64, "prefix slot size");
constexpr uint64_t PSLOT_MAX_H = 1ull << 28;

__global__ void k_prefix_slots(const uint32_t* __restrict__ kr, const uint32_t* __restrict__ pl,
                               const uint64_t* __restrict__ pk, uint64_t H, PrefixSlot* __restrict__ slots) {
    GRID_STRIDE(h, H) {
        PrefixSlot sl;
        sl.lo = kr[h];
        sl.cnt = kr[h + 1] - sl.lo;
#pragma unroll
        for (uint32_t k = 0; k < PSLOT_N; ++k) {
            sl.id[k] = k < sl.cnt ? pl[sl.lo + k] : INV;
            sl.key[k] = k < sl.cnt ? pk[sl.lo + k] : 0ull;
        }
        sl.pad[0] = sl.pad[1] = 0;
        slots[h] = sl;
    }
}

// a6 probe (compress.cu:1474-1528): hash the k-mer at kip+off, binary
// search the bucket, take the first free non-self candidate if it matches.
// With slots, buckets of up to PSLOT_N entries are searched in registers
// (the same lower bound and the same first-free scan).
// n bases from base `off` of suffix i, its words held word-major
// (word w of suffix i at sw[w * stride + i]): neighbouring threads read
// neighbouring words.
__device__ __forceinline__ uint64_t xw_soa(const uint64_t* __restrict__ sw, uint64_t stride, uint64_t i,
                                           uint32_t off, uint32_t n) {
    const uint32_t o1 = off >> 5;
    const uint32_t o2 = (off & 31u) << 1;
    const uint64_t left = sw[o1 * stride + i] << o2;
    if ((64u - o2) >= (n << 1)) return left >> ((32u - n) << 1);
    const uint64_t right = (sw[(o1 + 1) * stride + i] >> (62u - o2)) >> 2;
    return (left | right) >> ((32u - n) << 1);
}

// Word-major copy of the suffix list's reads.
__global__ void k_gather_soa(const uint64_t* __restrict__ R, const uint32_t* __restrict__ ids, uint64_t n,
                             uint32_t W, uint64_t* __restrict__ sw) {
    GRID_STRIDE(t, n * W) {
        const uint64_t w = t / n, i = t - w * n;
        sw[t] = R[(uint64_t)ids[i] * W + w];
    }
}

// Candidate step of the a6 probe for suffix index i whose k-mer hashed to
// bucket h: lower bound of s[off:L] in the bucket, then the first entry of
// the equal run that is neither s nor taken.  Sets *grp to the run start
// (the group id) and returns the chosen prefix position or INV.  With
// slots, buckets of up to PSLOT_N entries are searched in registers (the
// same lower bound and the same first-free scan).
__device__ __forceinline__ uint32_t ov_candidate(const ClaimState& cs, const uint32_t* __restrict__ kr,
                                                 const PrefixSlot* __restrict__ slots,
                                                 const uint64_t* __restrict__ sw, uint32_t ns, uint32_t i,
                                                 uint32_t h, uint32_t off, uint32_t* grp) {
    const uint32_t len = cs.L - off;
    const uint32_t n1 = min(32u, len);
    const uint32_t s = cs.sl[i];
    const uint64_t* rs = cs.R + (uint64_t)s * cs.W;
    // the suffix's words: from the word-major copy when there is one, else its packed row
    const uint64_t sx0 = sw ? xw_soa(sw, ns, i, off, n1) : xw(rs, off, n1);
    uint32_t lo, e;
    if (slots) {
        const PrefixSlot sl = slots[h];
        lo = sl.lo;
        e = lo + sl.cnt;
        if (sl.cnt <= PSLOT_N) {
            // lower bound: first entry not below s[off:L]
            uint32_t b = sl.cnt;
            bool found = false;
#pragma unroll
            for (uint32_t t = 0; t < PSLOT_N; ++t) {
                if (!found && t < sl.cnt && cmp_ps<Enc2>(cs.R, cs.W, sl.id[t], sl.key[t], rs, sx0, off, len) >= 0) {
                    b = t;
                    found = true;
                }
            }
            // first entry from there that is neither s nor taken decides;
            // entries are sorted, so past the equal run nothing can match
            // and the taken flags need not be read
            uint32_t ch = INV;
            bool done = false;
#pragma unroll
            for (uint32_t t = 0; t < PSLOT_N; ++t) {
                if (!done && t >= b && t < sl.cnt) {
                    const uint32_t p = sl.id[t];
                    if (cmp_ps<Enc2>(cs.R, cs.W, p, sl.key[t], rs, sx0, off, len) != 0) {
                        done = true;
                    } else if (p != s && cs.prev[p] == INV) {
                        ch = lo + t;
                        done = true;
                    }
                }
            }
            *grp = lo + b;
            return ch;
        }
    } else {
        lo = kr[h];
        e = kr[h + 1];
    }
    uint32_t cnt = e - lo;
    while (cnt > 0) {
        uint32_t step = cnt >> 1, it = lo + step;
        if (cmp_ps<Enc2>(cs.R, cs.W, cs.pl[it], cs.pk[it], rs, sx0, off, len) < 0) {
            lo = it + 1;
            cnt -= step + 1;
        } else {
            cnt = step;
        }
    }
    // the first entry that is neither s nor taken decides; past the run of
    // entries equal to s[off:L] nothing can match (sorted bucket)
    uint32_t ch = INV;
    for (uint32_t x = lo; x < e; ++x) {
        uint32_t p = cs.pl[x];
        if (cmp_ps<Enc2>(cs.R, cs.W, p, cs.pk[x], rs, sx0, off, len) != 0) break;
        if (p == s || cs.prev[p] != INV) continue;
        ch = x;
        break;
    }
    *grp = lo;
    return ch;
}

// ov_candidate for the verdict form, with rows held in registers.  The
// suffix's row is read once (its words requested together) and supplies the
// k-mer and every chunk of s[off:L); a bucket entry whose first word equals
// s[off:off+32) has its whole row requested at once and compared in
// registers.  Comparing from memory costs one more round of dependent loads
// per further 32-base chunk, and the lower bound and the first-free scan
// compare the same entry twice.  Same lower bound, same first-free scan,
// same result as ov_candidate; buckets of more than PSLOT_N entries (and
// indexes without slots) go to ov_candidate.
template <uint32_t WM>
__device__ __forceinline__ uint32_t ov_candidate_r(const ClaimState& cs, const uint32_t* __restrict__ kr,
                                                   const PrefixSlot* __restrict__ slots, uint32_t ns, uint32_t i,
                                                   uint32_t k, uint32_t kip, uint32_t mask, uint32_t off,
                                                   uint32_t* grp) {
    const uint32_t W = cs.W;
    const uint32_t len = cs.L - off;
    const uint32_t s = cs.sl[i];
    uint64_t r[WM];
    {
        const uint64_t* rs = cs.R + (uint64_t)s * W;
#pragma unroll
        for (uint32_t w = 0; w < WM; ++w) r[w] = w < W ? ld64(rs + w) : 0ull;
    }
    const uint32_t h = (uint32_t)(murmur64(xw_pick(r, kip + off, k)) & mask);
    if (!slots) return ov_candidate(cs, kr, slots, nullptr, ns, i, h, off, grp);
    const PrefixSlot sl = slots[h];
    if (sl.cnt > PSLOT_N) return ov_candidate(cs, kr, slots, nullptr, ns, i, h, off, grp);
    const uint32_t n1 = min(32u, len);
    const uint64_t sx0 = xw_pick(r, off, n1);
    // sign of p[0:len) - s[off:L) for each entry (its first word decides
    // unless it equals s[off:off+32))
    int res[PSLOT_N];
#pragma unroll
    for (uint32_t t = 0; t < PSLOT_N; ++t) {
        const uint64_t a = Enc2::head(sl.key[t], n1);
        res[t] = a == sx0 ? 0 : (a < sx0 ? -1 : 1);
    }
    if (len > 32) {
#pragma unroll
        for (uint32_t t = 0; t < PSLOT_N; ++t) {
            if (t < sl.cnt && res[t] == 0) {
                uint64_t q[WM];
                const uint64_t* rp = cs.R + (uint64_t)sl.id[t] * W;
#pragma unroll
                for (uint32_t w = 0; w < WM; ++w) q[w] = w < W ? ld64(rp + w) : 0ull;
                for (uint32_t j = 32; j < len; j += 32) {
                    const uint32_t n = min(32u, len - j);
                    const uint64_t x = xw_pick(q, j, n), y = xw_pick(r, off + j, n);
                    if (x != y) {
                        res[t] = x < y ? -1 : 1;
                        break;
                    }
                }
            }
        }
    }
    // lower bound: the first entry not below s[off:L); from there the first
    // entry that is neither s nor taken decides, and past the equal run
    // nothing can match (sorted bucket)
    uint32_t b = sl.cnt;
#pragma unroll
    for (uint32_t t = PSLOT_N; t-- > 0;)
        if (t < sl.cnt && res[t] >= 0) b = t;
    uint32_t ch = INV;
    bool done = false;
#pragma unroll
    for (uint32_t t = 0; t < PSLOT_N; ++t) {
        if (!done && t >= b && t < sl.cnt) {
            if (res[t] != 0) {
                done = true;
            } else {
                const uint32_t p = sl.id[t];
                if (p != s && cs.prev[p] == INV) {
                    ch = sl.lo + t;
                    done = true;
                }
            }
        }
    }
    *grp = sl.lo + b;
    return ch;
}

// k-mer filter verdict: a clear bit proves that no prefix has this k-mer
// (exact, no false negatives).
__device__ __forceinline__ bool ov_filter_pass(const uint32_t* __restrict__ filt, uint32_t fmask, uint64_t m) {
    if (!filt) return true;
    const uint32_t f = (uint32_t)(m >> 32) & fmask;
    return ((filt[f >> 5] >> (f & 31u)) & 1u) != 0;
}

// The suffix list is not compacted between offsets: a suffix linked at an
// earlier offset keeps its slinked flag and is skipped, and the suffixes'
// words come from the word-major copy sw (stride = list length).
__global__ void __launch_bounds__(256) k_ov_probe(ClaimState cs, const uint32_t* __restrict__ kr,
                                                  const PrefixSlot* __restrict__ slots,
                                                  const uint32_t* __restrict__ filt, uint32_t fmask, uint32_t mask,
                                                  const uint64_t* __restrict__ sw, uint32_t k, uint32_t kip,
                                                  uint32_t off) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *cs.res_n = *cs.pool_n = 0;
    const uint32_t ns = *cs.d_ns;
    GRID_STRIDE(i, ns) {
        if (cs.slinked[i]) {
            cs.choice[i] = INV;
            continue;
        }
        const uint64_t m = murmur64(xw_soa(sw, ns, i, kip + off, k));
        if (!ov_filter_pass(filt, fmask, m)) {
            cs.choice[i] = INV;
            continue;
        }
        uint32_t g;
        const uint32_t ch = ov_candidate(cs, kr, slots, sw, ns, (uint32_t)i, (uint32_t)(m & mask), off, &g);
        cs.choice[i] = ch;
        cs.grp[i] = g;
        if (ch != INV) {
            uint32_t old = atomicCAS(&cs.owner[ch], INV, (uint32_t)i);
            if (old != INV) cs.contended[g] = 1;
        }
    }
}

// Queued form of the a6 probe.  About one suffix in fifteen passes the
// k-mer filter, so in the plain form nearly every wave carries a few lanes
// through the candidate step's chain of dependent loads (slot, read words,
// taken flag, owner).  Here workgroup b owns the contiguous chunk
// [b*per, b*per + per) of the suffix list: it filters OVQ_TILE suffixes at
// a time (OVQ_ITEMS per thread, their loads issued together), queues the
// passing ones in an LDS ring, and runs the candidate step on full batches
// of 256, one per thread.  Suffixes that propose are written to the chunk's
// own segment of cand (cand[b*per ...], ncand[b] of them), which is all the
// finalisation needs; choice and grp are written for proposers only.  The
// proposals are the same as the plain form's: each depends on the suffix
// alone, and the claim outcome does not depend on proposal order.
constexpr uint32_t OVQ_ITEMS = 4;
constexpr uint32_t OVQ_TILE = 256 * OVQ_ITEMS;
constexpr uint32_t OVQ_CAP = 2048;  // ring: one tile plus a partial batch fit
static_assert(OVQ_CAP >= OVQ_TILE + 512 && (OVQ_CAP & (OVQ_CAP - 1)) == 0, "overlap queue capacity");

// VERD: the filter verdicts were computed for all offsets beforehand
// (k_ov_verdicts; bit i of vb[(off - 1) * nw + i / 64]), so the filter pass
// reads one bit per suffix and the candidate step (ov_candidate_r, rows of
// at most WM words) hashes the k-mer itself, from the suffix's packed row
// (sw is null: only the few suffixes that pass read words, and they read
// that row anyway).
template <bool VERD, uint32_t WM>
__global__ void __launch_bounds__(256) k_ov_probe_q(ClaimState cs, const uint32_t* __restrict__ kr,
                                                    const PrefixSlot* __restrict__ slots,
                                                    const uint32_t* __restrict__ filt, uint32_t fmask, uint32_t mask,
                                                    const uint64_t* __restrict__ sw, uint32_t k, uint32_t kip,
                                                    uint32_t off, uint32_t per, uint32_t* __restrict__ cand,
                                                    uint32_t* __restrict__ ncand, const uint64_t* __restrict__ vb,
                                                    uint64_t nw, uint32_t* __restrict__ passed) {
    __shared__ uint32_t qi[OVQ_CAP];
    __shared__ uint32_t qh[OVQ_CAP];
    __shared__ uint32_t qtail, nout;
    if (blockIdx.x == 0 && threadIdx.x == 0) *cs.res_n = *cs.pool_n = 0;
    const uint32_t ns = *cs.d_ns;
    const uint64_t beg64 = (uint64_t)blockIdx.x * per;
    const uint32_t beg = (uint32_t)min<uint64_t>(beg64, ns);
    const uint32_t end = (uint32_t)min<uint64_t>(beg64 + per, ns);
    const uint32_t lane = __lane_id();
    const uint64_t lt = lanemask_lt();
    if (threadIdx.x == 0) qtail = nout = 0;
    __syncthreads();

    // candidate step for queue entry q (valid: the thread holds an entry)
    auto candidate = [&](bool valid, uint32_t q) {
        uint32_t ch = INV, i = 0;
        if (valid) {
            i = qi[q & (OVQ_CAP - 1)];
            uint32_t g;
            if constexpr (VERD)
                ch = ov_candidate_r<WM>(cs, kr, slots, ns, i, k, kip, mask, off, &g);
            else
                ch = ov_candidate(cs, kr, slots, sw, ns, i, qh[q & (OVQ_CAP - 1)], off, &g);
            if (ch != INV) {
                cs.choice[i] = ch;
                cs.grp[i] = g;
                uint32_t old = atomicCAS(&cs.owner[ch], INV, i);
                if (old != INV) cs.contended[g] = 1;
            }
        }
        const uint64_t bal = __ballot(ch != INV);
        if (bal) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&nout, (uint32_t)__popcll(bal));
            base = __shfl(base, 0);
            if (ch != INV) cand[beg + base + (uint32_t)__popcll(bal & lt)] = i;
        }
    };

    uint32_t head = 0;
    for (uint32_t base = beg; base < end; base += OVQ_TILE) {
        uint32_t idx[OVQ_ITEMS];
        bool live[OVQ_ITEMS];
        uint64_t kw[OVQ_ITEMS];
#pragma unroll
        for (uint32_t j = 0; j < OVQ_ITEMS; ++j) {
            idx[j] = base + j * 256 + threadIdx.x;
            if (VERD)
                live[j] = idx[j] < end && ((vb[(uint64_t)(off - 1) * nw + (idx[j] >> 6)] >> (idx[j] & 63u)) & 1ull) &&
                          cs.slinked[idx[j]] == 0;
            else
                live[j] = idx[j] < end && cs.slinked[idx[j]] == 0;
        }
        if (!VERD) {
#pragma unroll
            for (uint32_t j = 0; j < OVQ_ITEMS; ++j) kw[j] = live[j] ? xw_soa(sw, ns, idx[j], kip + off, k) : 0ull;
        }
#pragma unroll
        for (uint32_t j = 0; j < OVQ_ITEMS; ++j) {
            const uint64_t m = VERD ? 0ull : murmur64(kw[j]);
            const bool pass = live[j] && (VERD || ov_filter_pass(filt, fmask, m));
            const uint64_t bal = __ballot(pass);
            if (bal) {
                uint32_t wb = 0;
                if (lane == 0) wb = atomicAdd(&qtail, (uint32_t)__popcll(bal));
                wb = __shfl(wb, 0);
                if (pass) {
                    const uint32_t q = (wb + (uint32_t)__popcll(bal & lt)) & (OVQ_CAP - 1);
                    qi[q] = idx[j];
                    if (!VERD) qh[q] = (uint32_t)(m & mask);
                }
            }
        }
        __syncthreads();
        const uint32_t t = qtail;
        while (t - head >= 256) {
            candidate(true, head + threadIdx.x);
            head += 256;
        }
        __syncthreads();
    }
    const uint32_t t = qtail;
    if (t > head) candidate(head + threadIdx.x < t, head + threadIdx.x);
    __syncthreads();
    if (threadIdx.x == 0) {
        ncand[blockIdx.x] = nout;
        // suffixes of this workgroup's segment that searched their bucket, summed over offsets
        if (passed) passed[blockIdx.x] += t;
    }
}

// Overlap probes of a static suffix list: a suffix linked at offset o was
// probed at offsets 1..o, an unlinked one at every offset below max_off.
__global__ void __launch_bounds__(256) k_ov_probe_count(const uint32_t* __restrict__ sl, uint32_t ns,
                                                        const uint16_t* __restrict__ off, uint32_t max_off,
                                                        uint64_t* __restrict__ out) {
    uint64_t acc = 0;
    GRID_STRIDE(i, ns) {
        const uint16_t o = off[sl[i]];
        acc += o == INVOFF ? (uint64_t)(max_off - 1) : (uint64_t)o;
    }
    // 64-bit totals per wave, then per workgroup: one atomic per workgroup
    // (same-address atomics are serialised, about 10 ns each)
    __shared__ uint64_t wsum[4];
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_down(acc, d, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint64_t t = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (t) atomicAdd((unsigned long long*)out, (unsigned long long)t);
    }
}

__global__ void __launch_bounds__(256) k_sum_u32(const uint32_t* __restrict__ a, uint32_t n,
                                                 uint64_t* __restrict__ out) {
    uint64_t acc = 0;
    GRID_STRIDE(i, n) acc += a[i];
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_down(acc, d, 64);
    if ((threadIdx.x & 63) == 0 && acc) atomicAdd((unsigned long long*)out, (unsigned long long)acc);
}

// Finalisation over the proposers k_ov_probe_q listed (same rules as
// claim_finalize_body): workgroup b walks its own segment.
__global__ void __launch_bounds__(256) k_ov_finalize_q(ClaimState cs, uint32_t off, uint32_t per,
                                                       const uint32_t* __restrict__ cand,
                                                       const uint32_t* __restrict__ ncand) {
    const uint32_t nc = ncand[blockIdx.x];
    const uint64_t beg = (uint64_t)blockIdx.x * per;
    for (uint32_t x = threadIdx.x; x < nc; x += blockDim.x) {
        const uint32_t i = cand[beg + x];
        const uint32_t ch = cs.choice[i];
        cs.owner[ch] = INV;
        const uint32_t g = cs.grp[i];
        const bool cont = cs.contended[g] != 0;
        // one atomic on the shared count per wavefront, not per member
        const uint32_t q = wave_append(cs.res_n, cont);
        if (cont) {
            cs.memb[i] = atomicExch(&cs.head[g], i);
            cs.res_list[q] = i;
        } else {
            cs.link(i, ch, off);
        }
    }
}

// ---- overlap probe by verdict words (tuning ov_queue 3) ------------------
// As the verdict form of k_ov_probe_q, but each thread takes 64 suffixes at
// once: the verdict word of the offset AND NOT the word of the linked-suffix
// bitmap (cs.slbits, set by ClaimState::link) leaves the suffixes that
// search their bucket, about one in twenty.  Workgroup b owns verdict words
// [b*wper, b*wper + wper) (suffixes b*wper*64 ...), OVW_TPW threads to a
// word (64 / OVW_TPW suffixes each), pushes their set bits into the LDS
// ring at most OVW_TAKE per thread and round, and runs the candidate step
// on full batches of 256, as k_ov_probe_q.  Its proposers go to its own
// segment of cand (per = wper * 64 entries), which k_ov_finalize_q walks.
constexpr uint32_t OVW_TAKE = 4;
constexpr uint32_t OVW_TPW = 4;
constexpr uint32_t OVW_TILE = 256 / OVW_TPW;  // words per round of the workgroup
static_assert(OVQ_CAP >= 256 * OVW_TAKE + 512, "overlap queue capacity");

template <uint32_t WM>
__global__ void __launch_bounds__(256) k_ov_probe_w(ClaimState cs, const uint32_t* __restrict__ kr,
                                                    const PrefixSlot* __restrict__ slots, uint32_t k, uint32_t kip,
                                                    uint32_t mask, uint32_t off, uint32_t wper,
                                                    const uint64_t* __restrict__ vb, uint64_t nw,
                                                    uint32_t* __restrict__ cand, uint32_t* __restrict__ ncand,
                                                    uint32_t* __restrict__ passed) {
    __shared__ uint32_t qi[OVQ_CAP];
    __shared__ uint32_t qtail, nout;
    if (blockIdx.x == 0 && threadIdx.x == 0) *cs.res_n = *cs.pool_n = 0;
    const uint32_t ns = *cs.d_ns;
    const uint64_t wb = (uint64_t)blockIdx.x * wper;
    const uint64_t we = min<uint64_t>(wb + wper, nw);
    const uint64_t beg = wb * 64;
    const uint32_t lane = __lane_id();
    const uint64_t lt = lanemask_lt();
    const uint64_t* vrow = vb + (uint64_t)(off - 1) * nw;
    if (threadIdx.x == 0) qtail = nout = 0;
    __syncthreads();

    auto candidate = [&](bool valid, uint32_t q) {
        uint32_t ch = INV, i = 0;
        if (valid) {
            i = qi[q & (OVQ_CAP - 1)];
            uint32_t g;
            ch = ov_candidate_r<WM>(cs, kr, slots, ns, i, k, kip, mask, off, &g);
            if (ch != INV) {
                cs.choice[i] = ch;
                cs.grp[i] = g;
                const uint32_t old = atomicCAS(&cs.owner[ch], INV, i);
                if (old != INV) cs.contended[g] = 1;
            }
        }
        const uint64_t bal = __ballot(ch != INV);
        if (bal) {
            uint32_t at = 0;
            if (lane == 0) at = atomicAdd(&nout, (uint32_t)__popcll(bal));
            at = __shfl(at, 0);
            if (ch != INV) cand[beg + at + (uint32_t)__popcll(bal & lt)] = i;
        }
    };

    uint32_t head = 0;
    constexpr uint32_t SB = 64 / OVW_TPW;  // suffixes per thread
    const uint32_t sub = threadIdx.x % OVW_TPW;
    for (uint64_t w0 = wb; w0 < we; w0 += OVW_TILE) {
        const uint64_t t = w0 + threadIdx.x / OVW_TPW;
        // verdict bits of suffixes past the list's end are clear
        uint64_t bits = t < we ? (vrow[t] & ~cs.slbits[t]) : 0ull;
        bits = (bits >> (SB * sub)) & (SB == 64 ? ~0ull : ((1ull << SB) - 1));
        const uint32_t i0 = (uint32_t)(t * 64 + SB * sub);
        while (__syncthreads_or(bits != 0)) {
            const uint32_t take = min((uint32_t)__popcll(bits), OVW_TAKE);
            uint32_t inc = take;
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t u = __shfl_up(inc, d, 64);
                if (lane >= (uint32_t)d) inc += u;
            }
            const uint32_t tot = __shfl(inc, 63);
            uint32_t qb = 0;
            if (lane == 0 && tot) qb = atomicAdd(&qtail, tot);
            uint32_t q = __shfl(qb, 0) + inc - take;
            for (uint32_t j = 0; j < take; ++j) {
                const uint32_t b = (uint32_t)__builtin_ctzll(bits);
                bits &= bits - 1;
                qi[q++ & (OVQ_CAP - 1)] = i0 + b;
            }
            __syncthreads();
            const uint32_t tl = qtail;
            while (tl - head >= 256) {
                candidate(true, head + threadIdx.x);
                head += 256;
            }
        }
    }
    __syncthreads();
    const uint32_t tl = qtail;
    if (tl > head) candidate(head + threadIdx.x < tl, head + threadIdx.x);
    __syncthreads();
    if (threadIdx.x == 0) {
        ncand[blockIdx.x] = nout;
        if (passed) passed[blockIdx.x] += tl;
    }
}

// a8 probe (compress.cu:1061-1160): the block of s[off:] (its first bp
// bases) bounds the search in the lexicographic prefix list; first equal
// non-self prefix.
template <class E>
__device__ __forceinline__ void fm_probe_body(const ClaimState& cs, const uint32_t* __restrict__ rng, uint32_t bp,
                                              uint32_t off) {
    const uint32_t ns = *cs.d_ns;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *cs.res_n = *cs.pool_n = 0;
        if (cs.work) {
            cs.work[0] += ns;
            if (off == 1) cs.work[1] = ns;
        }
    }
    const uint32_t len = cs.L - off;
    const uint32_t n1 = min(E::BPW, len);
    const uint32_t np = *cs.d_np;
    if (!rng) {
        // no block ranges (short lists): search the whole prefix list, and
        // clear the prefix flags the ranges kernel would have cleared
        GRID_STRIDE(i, np) cs.plinked[i] = 0;
    }
    GRID_STRIDE(i, ns) {
        cs.slinked[i] = 0;
        uint32_t s = cs.sl[i];
        const uint64_t* rs = cs.R + (uint64_t)s * cs.W;
        uint32_t lo = 0, e = np;
        if (rng) {
            const uint32_t b = (uint32_t)E::word(rs, off, bp);
            lo = rng[b];
            e = rng[b + 1];
        }
        uint64_t sx0 = E::word(rs, off, n1);
        uint32_t cnt = e - lo;
        while (cnt > 0) {
            uint32_t step = cnt >> 1, it = lo + step;
            if (cmp_ps<E>(cs.R, cs.W, cs.pl[it], cs.pk[it], rs, sx0, off, len) < 0) {
                lo = it + 1;
                cnt -= step + 1;
            } else {
                cnt = step;
            }
        }
        uint32_t ch = INV;
        if (lo < e && cmp_ps<E>(cs.R, cs.W, cs.pl[lo], cs.pk[lo], rs, sx0, off, len) == 0) {
            if (cs.pl[lo] != s) {
                ch = lo;
            } else if (lo + 1 < e && cmp_ps<E>(cs.R, cs.W, cs.pl[lo + 1], cs.pk[lo + 1], rs, sx0, off, len) == 0) {
                ch = lo + 1;
            }
        }
        cs.choice[i] = ch;
        cs.grp[i] = lo;
        if (ch != INV) {
            uint32_t old = atomicCAS(&cs.owner[ch], INV, (uint32_t)i);
            if (old != INV) cs.contended[lo] = 1;
        }
    }
}

template <class E>
__global__ void __launch_bounds__(256) k_fm_probe(ClaimState cs, const uint32_t* __restrict__ rng, uint32_t bp,
                                                  uint32_t off) {
    fm_probe_body<E>(cs, rng, bp, off);
}

// Uncontended groups link at once; members of contended groups are queued
// for their leader.
// zero_p / zero_n: a buffer the next round needs cleared (the block ranges
// of the exact-matching loop), folded in here to save a launch.
__device__ __forceinline__ void claim_finalize_body(const ClaimState& cs, uint32_t off, uint32_t* zero_p,
                                                    uint64_t zero_n) {
    GRID_STRIDE(z, zero_n) zero_p[z] = 0;
    const uint32_t ns = *cs.d_ns;
    GRID_STRIDE(i, ns) {
        uint32_t ch = cs.choice[i];
        if (ch == INV) continue;
        cs.owner[ch] = INV;
        uint32_t g = cs.grp[i];
        if (cs.contended[g]) {
            cs.memb[i] = atomicExch(&cs.head[g], (uint32_t)i);
            uint32_t q = atomicAdd(cs.res_n, 1u);
            cs.res_list[q] = (uint32_t)i;
        } else {
            cs.link((uint32_t)i, ch, off);
        }
    }
}

__global__ void __launch_bounds__(256) k_claim_finalize(ClaimState cs, uint32_t off, uint32_t* zero_p = nullptr,
                                                        uint64_t zero_n = 0) {
    claim_finalize_body(cs, off, zero_p, zero_n);
}

// Suffix order inside a group: a6 keeps the suffix list order (compress.cu:
// 1474, one thread per list entry, sequential meaning); a8 orders by the
// reversed s[0:off) (see cmp_rev_prefix).
template <bool FULLMATCH, class E>
__device__ __forceinline__ bool member_before(const ClaimState& cs, uint32_t u, uint32_t v, uint32_t off) {
    if (FULLMATCH) return cmp_rev_prefix<E>(cs.R, cs.W, cs.sl[u], cs.sl[v], off) < 0;
    return u < v;
}

// First comparison word of cmp_rev_prefix: the n = min(BPW, off) bases
// before `off`, reversed.
template <class E>
__device__ __forceinline__ uint64_t rev_prefix_key(const uint64_t* __restrict__ R, uint32_t W, uint32_t s,
                                                   uint32_t off) {
    const uint32_t n = min(E::BPW, off);
    return E::reversed(E::word(R + (uint64_t)s * W, off - n, n), n);
}

// Members a[0..m) of a large a8 group in cmp_rev_prefix order, with each
// member's first comparison word loaded once into k[0..m) (heapsort on the
// pairs; the reads are touched again only to break ties of that word).
template <class E>
__device__ __forceinline__ void sort_members_keyed(const ClaimState& cs, uint32_t* a, uint64_t* k, uint32_t m, uint32_t off) {
    for (uint32_t i = 0; i < m; ++i) k[i] = rev_prefix_key<E>(cs.R, cs.W, cs.sl[a[i]], off);
    auto before = [&](uint32_t x, uint32_t y) {
        if (k[x] != k[y]) return k[x] < k[y];
        return cmp_rev_prefix<E>(cs.R, cs.W, cs.sl[a[x]], cs.sl[a[y]], off) < 0;
    };
    auto swap = [&](uint32_t x, uint32_t y) {
        const uint32_t t = a[x];
        a[x] = a[y];
        a[y] = t;
        const uint64_t u = k[x];
        k[x] = k[y];
        k[y] = u;
    };
    auto sift = [&](uint32_t root, uint32_t end) {
        for (;;) {
            uint32_t c = 2 * root + 1;
            if (c >= end) return;
            if (c + 1 < end && before(c, c + 1)) ++c;
            if (!before(root, c)) return;
            swap(root, c);
            root = c;
        }
    };
    for (uint32_t i = m / 2; i-- > 0;) sift(i, m);
    for (uint32_t end = m; end-- > 1;) {
        swap(0, end);
        sift(0, end);
    }
}

template <bool FULLMATCH, class E>
__device__ __forceinline__ void sort_members(const ClaimState& cs, uint32_t* a, uint32_t m, uint32_t off) {
    if (m <= 16) {
        for (uint32_t i = 1; i < m; ++i) {
            uint32_t x = a[i], j = i;
            while (j > 0 && member_before<FULLMATCH, E>(cs, x, a[j - 1], off)) {
                a[j] = a[j - 1];
                --j;
            }
            a[j] = x;
        }
        return;
    }
    // heapsort (members are distinct and strictly ordered)
    auto sift = [&](uint32_t root, uint32_t end) {
        for (;;) {
            uint32_t c = 2 * root + 1;
            if (c >= end) return;
            if (c + 1 < end && member_before<FULLMATCH, E>(cs, a[c], a[c + 1], off)) ++c;
            if (!member_before<FULLMATCH, E>(cs, a[root], a[c], off)) return;
            uint32_t t = a[root];
            a[root] = a[c];
            a[c] = t;
            root = c;
        }
    };
    for (uint32_t i = m / 2; i-- > 0;) sift(i, m);
    for (uint32_t end = m; end-- > 1;) {
        uint32_t t = a[0];
        a[0] = a[end];
        a[end] = t;
        sift(0, end);
    }
}

// First free prefix position >= x of a list kept in place (np: none).
// nf[x] == x for a free position; linking x sets nf[x] = x + 1, so every
// position in [x, nf[x]) is linked, and the walk shortens the path it takes
// (nf[x] = nf[nf[x]]).  Values only grow and each one is a valid bound at
// any time, so walks racing with each other or with links elsewhere stay
// correct: a stale value only makes a walk longer.
__device__ __forceinline__ uint32_t fm_find(uint32_t* __restrict__ nf, uint32_t x, uint32_t np) {
    while (x < np) {
        const uint32_t y = nf[x];
        if (y == x) return x;
        const uint32_t z = (y < np) ? nf[y] : y;
        if (z != y) nf[x] = z;
        x = z;
    }
    return np;
}

// One leader per contended group replays it as the reference's sequential
// order would: members sorted by suffix order, each taking the first free
// non-self prefix of the group's run (compress.cu:1500-1527 for a6; the
// merge join with self-avoid rotation, compress.cu:1089-1160, for a8).  The
// free part of the run is always {P[a]} followed by P[b..], so one pass
// suffices.
// The leader's walk over a sorted group arr[0..m) whose prefixes start at
// run position g.  All members share s[off:L], and the prefixes equal to it
// form one run [g, e) of the sorted list (one bucket's sorted run in a6),
// which starts at g.  Its end is found once, by galloping, instead of one
// comparison per linked member.
template <class E, typename At>
__device__ __forceinline__ void claim_assign_at(const ClaimState& cs, At arr, uint32_t m, uint32_t g, uint32_t off) {
    const uint32_t np = *cs.d_np;
    const uint32_t len = cs.L - off;
    const uint32_t n1 = min(E::BPW, len);
    const uint32_t s0 = cs.sl[arr(0)];
    const uint64_t* rs0 = cs.R + (uint64_t)s0 * cs.W;
    const uint64_t sx00 = E::word(rs0, off, n1);
    auto equal = [&](uint32_t x) { return cmp_ps<E>(cs.R, cs.W, cs.pl[x], cs.pk[x], rs0, sx00, off, len) == 0; };
    uint32_t known = g, e = g + 1;  // [g, known] equal; e: first candidate past it
    uint64_t step = 1;
    while (e < np && equal(e)) {
        known = e;
        step <<= 1;
        e = ((uint64_t)(np - g) > step) ? g + (uint32_t)step : np;
    }
    while (e - known > 1) {
        const uint32_t mid = known + (e - known) / 2;
        if (equal(mid))
            known = mid;
        else
            e = mid;
    }
    auto next_free = [&](uint32_t x) -> uint32_t {
        if (cs.nf) {
            const uint32_t f = fm_find(cs.nf, x, np);
            return f < e ? f : e;
        }
        while (x < e && cs.prev[cs.pl[x]] != INV) ++x;
        return x;
    };
    uint32_t a = next_free(g);
    uint32_t b = next_free(a + 1);
    for (uint32_t k = 0; k < m && a < e; ++k) {
        const uint32_t s = cs.sl[arr(k)];
        if (cs.pl[a] != s) {
            cs.link(arr(k), a, off);
            a = b;
            b = next_free(b + 1);
        } else if (b < e) {
            cs.link(arr(k), b, off);
            b = next_free(b + 1);
        }
    }
    cs.head[g] = INV;
    if (cs.contended) cs.contended[g] = 0;
}

template <class E>
__device__ __forceinline__ void claim_assign(const ClaimState& cs, const uint32_t* arr, uint32_t m, uint32_t g,
                                             uint32_t off) {
    claim_assign_at<E>(cs, [&](uint32_t k) { return arr[k]; }, m, g, off);
}

// a8 groups of RESOLVE_COOP_MIN < m <= RESOLVE_COOP_MAX members are sorted
// by the whole workgroup (bitonic sort in LDS on the same order as
// sort_members_keyed: first comparison word, then the reads) after its
// grid-stride pass, instead of by the leader alone; up to RESOLVE_COOP_LIST
// per workgroup and call, the rest as before.
constexpr uint32_t RESOLVE_COOP_MIN = 16;
constexpr uint32_t RESOLVE_COOP_MAX = 4096;
constexpr uint32_t RESOLVE_COOP_LIST = 32;
struct ResolveCoop {
    uint64_t key[RESOLVE_COOP_MAX];
    uint32_t mem[RESOLVE_COOP_MAX];
    uint32_t lead[RESOLVE_COOP_LIST], base[RESOLVE_COOP_LIST], m[RESOLVE_COOP_LIST];
    uint32_t n;
};

template <class E>
__device__ __forceinline__ void resolve_coop(const ClaimState& cs, ResolveCoop& co, uint32_t off) {
    __syncthreads();
    const uint32_t nb = min(co.n, RESOLVE_COOP_LIST);
    for (uint32_t gi = 0; gi < nb; ++gi) {
        const uint32_t m = co.m[gi];
        const uint32_t* arr = cs.pool + co.base[gi];
        uint32_t P = 1;
        while (P < m) P <<= 1;
        for (uint32_t x = threadIdx.x; x < P; x += blockDim.x) {
            if (x < m) {
                const uint32_t u = arr[x];
                co.mem[x] = u;
                co.key[x] = rev_prefix_key<E>(cs.R, cs.W, cs.sl[u], off);
            } else {
                co.mem[x] = INV;  // padding sorts last
                co.key[x] = ~0ull;
            }
        }
        __syncthreads();
        auto before = [&](uint32_t x, uint32_t y) -> bool {
            const uint32_t ux = co.mem[x], uy = co.mem[y];
            if (ux == INV) return false;
            if (uy == INV) return true;
            const uint64_t kx = co.key[x], ky = co.key[y];
            if (kx != ky) return kx < ky;
            return cmp_rev_prefix<E>(cs.R, cs.W, cs.sl[ux], cs.sl[uy], off) < 0;
        };
        for (uint32_t k = 2; k <= P; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t x = threadIdx.x; x < P; x += blockDim.x) {
                    const uint32_t y = x ^ j;
                    if (y > x && (((x & k) == 0) ? before(y, x) : before(x, y))) {
                        const uint32_t tm = co.mem[x];
                        co.mem[x] = co.mem[y];
                        co.mem[y] = tm;
                        const uint64_t tk = co.key[x];
                        co.key[x] = co.key[y];
                        co.key[y] = tk;
                    }
                }
                __syncthreads();
            }
        }
        if (threadIdx.x == 0) claim_assign<E>(cs, co.mem, m, cs.grp[co.lead[gi]], off);
        __syncthreads();
    }
}

template <bool FULLMATCH, class E>
__device__ __forceinline__ void claim_resolve_body(const ClaimState& cs, uint32_t off) {
    const uint32_t nq = *cs.res_n;
    ResolveCoop* co = nullptr;
    if constexpr (FULLMATCH) {
        __shared__ ResolveCoop coop;
        co = &coop;
        if (threadIdx.x == 0) co->n = 0;
        __syncthreads();
    }
    GRID_STRIDE(q, nq) {
        uint32_t lead = cs.res_list[q];
        uint32_t g = cs.grp[lead];
        if (cs.head[g] != lead) continue;
        uint32_t m = 0;
        for (uint32_t u = lead; u != INV; u = cs.memb[u]) ++m;
        const uint32_t base = atomicAdd(cs.pool_n, m);
        uint32_t* arr = cs.pool + base;
        m = 0;
        for (uint32_t u = lead; u != INV; u = cs.memb[u]) arr[m++] = u;
        if (FULLMATCH && cs.kpool && m > RESOLVE_COOP_MIN && m <= RESOLVE_COOP_MAX) {
            const uint32_t slot = atomicAdd(&co->n, 1u);
            if (slot < RESOLVE_COOP_LIST) {
                co->lead[slot] = lead;
                co->base[slot] = base;
                co->m[slot] = m;
                continue;
            }
        }
        if (FULLMATCH && cs.kpool && m > 16)
            sort_members_keyed<E>(cs, arr, cs.kpool + base, m, off);
        else
            sort_members<FULLMATCH, E>(cs, arr, m, off);
        claim_assign<E>(cs, arr, m, g, off);
    }
    if constexpr (FULLMATCH) {
        if (cs.kpool) resolve_coop<E>(cs, *co, off);
    }
}

template <bool FULLMATCH, class E>
__global__ void __launch_bounds__(256) k_claim_resolve(ClaimState cs, uint32_t off) {
    claim_resolve_body<FULLMATCH, E>(cs, off);
}

// Block ranges of the sorted prefix list in one launch: rng[b] is the first
// list position whose block value (first bp bases) is >= b, for b in
// [0, nb], with rng[nb] = np.  Position i writes the values in
// (block(i-1), block(i)] (position np: up to nb), so every entry is written
// exactly once and no histogram, scan or clearing is needed.  Gaps of 64
// or more values are written by the whole wave.  Also clears the prefix
// flags of the list (as the histogram kernel does).
template <class E>
__global__ void __launch_bounds__(256) k_fm_ranges(const uint64_t* __restrict__ pk, const uint32_t* __restrict__ d_np,
                                                   uint32_t bp, uint32_t nb, uint32_t* __restrict__ rng,
                                                   uint8_t* __restrict__ plinked) {
    const uint32_t np = *d_np;
    const uint32_t lane = __lane_id();
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base <= np; base += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = base + threadIdx.x;
        uint32_t lo = 1, hi = 0;  // empty unless this position starts a gap
        if (i <= np) {
            if (i < np && plinked) plinked[i] = 0;
            lo = (i == 0) ? 0u : (uint32_t)E::head(pk[i - 1], bp) + 1u;
            hi = (i == np) ? nb : (uint32_t)E::head(pk[i], bp);
        }
        const bool have = i <= np && lo <= hi;
        if (have && hi - lo < 64)
            for (uint32_t b = lo; b <= hi; ++b) rng[b] = (uint32_t)i;
        uint64_t wide = __ballot(have && hi - lo >= 64);
        while (wide) {
            const int l = __ffsll((long long)wide) - 1;
            wide &= wide - 1;
            const uint32_t wlo = __shfl(lo, l), whi = __shfl(hi, l), v = __shfl((uint32_t)i, l);
            for (uint32_t b = wlo + lane; b <= whi; b += 64) rng[b] = v;
        }
    }
}

struct PredSufFree {
    const uint32_t* sl;
    const uint32_t* next;
    __device__ bool operator()(uint32_t i) const { return next[sl[i]] == INV; }
};
struct PredPreFree {
    const uint32_t* pl;
    const uint32_t* prev;
    __device__ bool operator()(uint32_t i) const { return prev[pl[i]] == INV; }
};
struct EmitPre {
    const uint32_t* pl;
    const uint64_t* pk;
    uint32_t* pl2;
    uint64_t* pk2;
    __device__ void operator()(uint32_t i, uint32_t j) const {
        pl2[j] = pl[i];
        pk2[j] = pk[i];
    }
};

// The whole exact-matching offset loop for short lists in one launch of one
// workgroup: per offset the probe, the claim finalisation and resolution
// and the two stable compactions run as phases separated by workgroup
// barriers (the same device code as the per-offset kernels, whose
// grid-stride loops cover the lists with this workgroup's threads).  c0 / c1
// are the claim states of the ping-pong list pairs; the loop ends early once
// either list is empty.
template <class E>
__global__ void __launch_bounds__(1024) k_fm_fused(ClaimState c0, ClaimState c1, uint32_t L) {
    int cur = 0;
    for (uint32_t off = 1; off < L; ++off) {
        const ClaimState& cs = cur ? c1 : c0;
        const ClaimState& ct = cur ? c0 : c1;
        fm_probe_body<E>(cs, nullptr, 0, off);
        __syncthreads();
        claim_finalize_body(cs, off, nullptr, 0);
        __syncthreads();
        claim_resolve_body<true, E>(cs, off);
        __syncthreads();
        compact_small_body(cs.d_ns, PredFlagClear{cs.slinked}, EmitGather{cs.sl, const_cast<uint32_t*>(ct.sl)},
                           const_cast<uint32_t*>(ct.d_ns));
        compact_small_body(cs.d_np, PredFlagClear{cs.plinked},
                           EmitPre{cs.pl, cs.pk, const_cast<uint32_t*>(ct.pl), const_cast<uint64_t*>(ct.pk)},
                           const_cast<uint32_t*>(ct.d_np));
        __syncthreads();
        cur ^= 1;
        if (*ct.d_ns == 0 || *ct.d_np == 0) break;
    }
}

// ------------------------------------------------ exact matching, groups ---
constexpr uint32_t FM_SMALL = 16;  // groups up to this size are sorted in registers
// Row width of the group table gmem (member slots per prefix position).
// Over lists kept in place it is FM_SMALL.  Over compacted lists (gcap: the
// table's entries) it is the largest power of two up to FM_ROW_MAX whose
// rows for all np positions fit, so that the larger groups of the last
// offsets (short lists) have their members listed too; the probe and the
// resolution derive it alike from the same counter.
constexpr uint32_t FM_ROW_MAX = 4096;
constexpr uint32_t FM_WAVE_MAX = 62;  // groups one wavefront resolves (pos(0 .. m+1), a lane each)
__device__ __forceinline__ uint32_t fm_row_width(uint64_t gcap, uint32_t np) {
    uint32_t w = FM_SMALL;
    while (w < FM_ROW_MAX && (uint64_t)np * (2 * w) <= gcap) w <<= 1;
    return w;
}
constexpr uint32_t FM_ROW = 0xFFFFFFFEu;  // FmCoop::base of a group whose members are in its row
// The exact-matching loop of k_fm_probe_s / k_fm_resolve_s, over lists kept
// in place (cs.nf set) or compacted after every offset (cs.nf null).
//
// Lists in place: the reference drops linked reads from both lists after
// every offset (compress.cu:1162-1172); here a linked suffix is skipped by
// its flag (slinked, kept set) and a linked prefix by the skip pointers nf
// (fm_find).  The free entries keep their order, so every suffix meets the
// same free prefixes in the same order as over the compacted lists, and the
// block ranges of the prefix list change only with the block width.  A
// suffix whose filter verdict at this offset is clear (vb: no prefix starts
// with its 32 bases from off) is not searched.
//
// Each suffix with an equal free non-self prefix joins its group (the first
// equal position of the list, grp) and records that prefix (choice): it
// takes the next rank of the group's count gcnt, goes into the group's row
// of gmem when the rank is below the row width (fm_row_width; so that the
// resolution reads a small group with one load), and into the group's
// member list (head / memb) for larger ones.  Rank 0 sets its lead flag, by which the
// resolution finds the group.  No counter is shared by all suffixes.
// zero_p / zero_n: a buffer to clear on the way (the look-back status words
// of the compaction that follows, between compacted offsets).
template <class E>
__global__ void __launch_bounds__(256) k_fm_probe_s(ClaimState cs, const uint32_t* __restrict__ rng, uint32_t bp,
                                                    uint32_t off, const uint64_t* __restrict__ vb, uint64_t vnw,
                                                    uint8_t* __restrict__ lead, uint32_t* __restrict__ gcnt,
                                                    uint32_t* __restrict__ gmem, uint64_t gcap,
                                                    uint32_t* __restrict__ zero_p, uint64_t zero_n) {
    const uint32_t ns = *cs.d_ns, np = *cs.d_np;
    const uint32_t rw = fm_row_width(gcap, np);
    const bool inplace = cs.nf != nullptr;
    // a buffer the next compaction needs cleared (its look-back status)
    GRID_STRIDE(z, zero_n) zero_p[z] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *cs.pool_n = 0;
        if (cs.work) {
            if (!inplace) cs.work[0] += ns;
            if (off == 1) cs.work[1] = ns;
        }
    }
    // compacted lists without block ranges: clear the prefix flags here
    if (!inplace && !rng) GRID_STRIDE(i, np) cs.plinked[i] = 0;
    const uint32_t len = cs.L - off;
    const uint32_t n1 = min(E::BPW, len);
    const uint64_t* vrow = vb ? vb + (uint64_t)(off - 1) * vnw : nullptr;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < ns; base += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = base + threadIdx.x;
        bool act = i < ns;
        if (act) {
            // compacted: every listed suffix is unlinked (its flag is of the last offset)
            if (inplace)
                act = cs.slinked[i] == 0;
            else
                cs.slinked[i] = 0;
        }
        const bool pr = act && (!vrow || ((vrow[i >> 6] >> (i & 63u)) & 1ull));
        if (!pr) continue;
        const uint32_t s = cs.sl[i];
        const uint64_t* rs = cs.R + (uint64_t)s * cs.W;
        uint32_t lo = 0, e = np;
        if (rng) {
            const uint32_t b = (uint32_t)E::word(rs, off, bp);
            lo = rng[b];
            e = rng[b + 1];
        }
        const uint64_t sx0 = E::word(rs, off, n1);
        uint32_t cnt = e - lo;
        while (cnt > 0) {
            const uint32_t step = cnt >> 1, it = lo + step;
            if (cmp_ps<E>(cs.R, cs.W, cs.pl[it], cs.pk[it], rs, sx0, off, len) < 0) {
                lo = it + 1;
                cnt -= step + 1;
            } else {
                cnt = step;
            }
        }
        // lo: the first position >= s[off:L] (the group id when equal)
        const uint32_t a = inplace ? fm_find(cs.nf, lo, np) : lo;
        uint32_t ch = INV;
        if (a < e && cmp_ps<E>(cs.R, cs.W, cs.pl[a], cs.pk[a], rs, sx0, off, len) == 0) {
            if (cs.pl[a] != s) {
                ch = a;
            } else {
                const uint32_t b2 = inplace ? fm_find(cs.nf, a + 1, np) : a + 1;
                if (b2 < e && cmp_ps<E>(cs.R, cs.W, cs.pl[b2], cs.pk[b2], rs, sx0, off, len) == 0) ch = b2;
            }
        }
        if (ch == INV) continue;
        cs.choice[i] = ch;
        cs.grp[i] = lo;
        const uint32_t r = atomicAdd(&gcnt[lo], 1u);
        if (r < rw) gmem[(uint64_t)lo * rw + r] = (uint32_t)i;
        cs.memb[i] = atomicExch(&cs.head[lo], (uint32_t)i);
        if (r == 0) lead[i] = 1;
    }
}

// Work of the offsets [1, Ls) of k_fm_probe_s, counted once afterwards: a
// suffix linked at offset o was considered at offsets 1..o, an unlinked one
// at 1..Ls-1; it was searched at those of them with a set verdict bit, and
// at every one from vmax on (no verdicts).  Adds the totals to out[0]
// (considered) and out[2] (searched), one atomic per workgroup.
__global__ void __launch_bounds__(256) k_fm_count_work(const uint32_t* __restrict__ sl, const uint32_t* __restrict__ d_ns,
                                                       const uint32_t* __restrict__ next,
                                                       const uint16_t* __restrict__ offset, uint32_t Ls,
                                                       const uint64_t* __restrict__ vb, uint64_t vnw, uint32_t vmax,
                                                       uint64_t* __restrict__ out) {
    __shared__ uint64_t part[2][4];
    const uint32_t ns = *d_ns;
    uint64_t act = 0, srch = 0;
    GRID_STRIDE(i, ns) {
        const uint32_t s = sl[i];
        const uint32_t o = next[s] != INV ? (uint32_t)offset[s] : Ls - 1;
        act += o;
        uint32_t nv = 0;  // offsets 1 .. min(o, vmax - 1) with a verdict
        if (vb) {
            const uint32_t top = min(o, vmax - 1);
            for (uint32_t f = 1; f <= top; ++f) nv += (uint32_t)((vb[(uint64_t)(f - 1) * vnw + (i >> 6)] >> (i & 63)) & 1ull);
            srch += nv + (o > top ? o - top : 0u);
        } else {
            srch += o;
        }
    }
    for (int d = 32; d >= 1; d >>= 1) {
        act += __shfl_down(act, d, 64);
        srch += __shfl_down(srch, d, 64);
    }
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[0][w] = act;
        part[1][w] = srch;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint64_t a = part[0][0] + part[0][1] + part[0][2] + part[0][3];
        const uint64_t b = part[1][0] + part[1][1] + part[1][2] + part[1][3];
        if (a) atomicAdd((unsigned long long*)&out[0], (unsigned long long)a);
        if (b) atomicAdd((unsigned long long*)&out[2], (unsigned long long)b);
    }
}

// cs.link with the suffix and prefix read ids already known.
__device__ __forceinline__ void fm_link(const ClaimState& cs, uint32_t si, uint32_t s, uint32_t px, uint32_t p,
                                        uint32_t off) {
    cs.next[s] = p;
    cs.prev[p] = s;
    cs.offset[s] = (uint16_t)off;
    cs.slinked[si] = 1;
    if (cs.plinked) cs.plinked[px] = 1;
    if (cs.nf) cs.nf[px] = px + 1;
}

// Member order of a group (the order of member_before / cmp_rev_prefix):
// first reversed-prefix word, then the reads; INV entries last.
template <class E>
__device__ __forceinline__ bool fm_before(const ClaimState& cs, uint32_t ux, uint64_t kx, uint32_t sx, uint32_t uy,
                                          uint64_t ky, uint32_t sy, uint32_t off) {
    if (ux == INV) return false;
    if (uy == INV) return true;
    if (kx != ky) return kx < ky;
    return cmp_rev_prefix<E>(cs.R, cs.W, sx, sy, off) < 0;
}

// The walk of claim_assign over a run whose positions are all free (lists
// compacted after every offset: the run [g, e) holds no linked prefix, and
// only this group links into it), with the members' reads sid[0..m) in
// sorted order and the run's first positions known: pos(j) is the prefix
// read at g + j, or INV past the run, for j <= m + 1 (the walk never looks
// further: b grows by one per linked member).  a / b are the two free
// positions the walk holds (claim_assign's a and b).
template <typename SidAt, typename MemAt, typename PosAt>
__device__ __forceinline__ void fm_walk_free(const ClaimState& cs, uint32_t m, uint32_t g, uint32_t off, SidAt sid,
                                             MemAt mem, PosAt pos) {
    uint32_t a = 0, b = 1;
    for (uint32_t k = 0; k < m; ++k) {
        const uint32_t pa = pos(a);
        if (pa == INV) break;
        const uint32_t s = sid(k);
        if (pa != s) {
            fm_link(cs, mem(k), s, g + a, pa, off);
            a = b;
            b = b + 1;
        } else {
            const uint32_t pb = pos(b);
            if (pb != INV) {
                fm_link(cs, mem(k), s, g + b, pb, off);
                b = b + 1;
            }
        }
    }
}

constexpr uint32_t FM_GATHER_MAX = 1u << 15;  // compacted lists up to this length: large groups gathered by scan

// The members' group marks: grp[i] is INV for every suffix position except
// those matched at the current offset (set by k_fm_probe_s); resolving a
// group clears its members' marks, so that a workgroup can gather a
// group's members by scanning the marks.

// A group walked by one thread over the pool: the members in list order
// into the pool, sorted (keyed heapsort), walked, and their marks cleared.
template <class E>
__device__ __forceinline__ void fm_serial_group(const ClaimState& cs, uint32_t g, uint32_t off) {
    const uint32_t u0 = cs.head[g];
    uint32_t m = 0;
    for (uint32_t u = u0; u != INV; u = cs.memb[u]) ++m;
    const uint32_t base = atomicAdd(cs.pool_n, m);
    uint32_t* arr = cs.pool + base;
    m = 0;
    for (uint32_t u = u0; u != INV; u = cs.memb[u]) arr[m++] = u;
    sort_members_keyed<E>(cs, arr, cs.kpool + base, m, off);
    claim_assign<E>(cs, arr, m, g, off);
    for (uint32_t k = 0; k < m; ++k) cs.grp[arr[k]] = INV;
}

// A group of 2..FM_SMALL members (mem[0..m), suffix positions; the rest
// INV): sorted by a bitonic network in registers, then walked.
template <class E>
__device__ __forceinline__ void fm_small_group(const ClaimState& cs, uint32_t (&mem)[FM_SMALL], uint32_t m, uint32_t g,
                                               uint32_t off) {
    uint64_t key[FM_SMALL];
    uint32_t sid[FM_SMALL];
#pragma unroll
    for (uint32_t k = 0; k < FM_SMALL; ++k) {
        sid[k] = mem[k] != INV ? cs.sl[mem[k]] : INV;
        key[k] = mem[k] != INV ? rev_prefix_key<E>(cs.R, cs.W, sid[k], off) : ~0ull;
    }
#pragma unroll
    for (uint32_t k = 2; k <= FM_SMALL; k <<= 1) {
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (uint32_t x = 0; x < FM_SMALL; ++x) {
                const uint32_t y = x ^ j;
                if (y > x) {
                    // ascending where x & k is clear: y must not come before x
                    const bool up = (x & k) == 0;
                    const bool sw = up ? fm_before<E>(cs, mem[y], key[y], sid[y], mem[x], key[x], sid[x], off)
                                       : fm_before<E>(cs, mem[x], key[x], sid[x], mem[y], key[y], sid[y], off);
                    if (sw) {
                        const uint32_t tm = mem[x], ts = sid[x];
                        const uint64_t tk = key[x];
                        mem[x] = mem[y];
                        sid[x] = sid[y];
                        key[x] = key[y];
                        mem[y] = tm;
                        sid[y] = ts;
                        key[y] = tk;
                    }
                }
            }
        }
    }
    if (cs.nf) {
        // lists in place (free positions by skip pointer): the member list
        // relinked in sorted order, and claim_assign walks it
#pragma unroll
        for (uint32_t k = 0; k + 1 < FM_SMALL; ++k)
            if (k + 1 < m) cs.memb[mem[k]] = mem[k + 1];
        uint32_t at_k = 0, at_u = mem[0];
        claim_assign_at<E>(
            cs,
            [&](uint32_t k) {
                while (at_k < k) {
                    at_u = cs.memb[at_u];
                    ++at_k;
                }
                return at_u;
            },
            m, g, off);
    } else {
        // compacted lists: fm_walk_free's walk over the run's first m + 2
        // positions, with the members in registers
        const uint32_t np = *cs.d_np, len = cs.L - off, n1 = min(E::BPW, len);
        const uint64_t* rs0 = cs.R + (uint64_t)sid[0] * cs.W;
        const uint64_t sx00 = E::word(rs0, off, n1);
        // the comparisons are independent (their loads overlap); the run is
        // the positions before the first unequal one
        uint32_t pr[FM_SMALL + 2];
#pragma unroll
        for (uint32_t j = 0; j < FM_SMALL + 2; ++j) {
            const uint32_t x = g + j;
            const bool eq = j <= m + 1 && x < np && cmp_ps<E>(cs.R, cs.W, cs.pl[x], cs.pk[x], rs0, sx00, off, len) == 0;
            pr[j] = eq ? cs.pl[x] : INV;
        }
        bool in = true;
#pragma unroll
        for (uint32_t j = 0; j < FM_SMALL + 2; ++j) {
            in = in && pr[j] != INV;
            if (!in) pr[j] = INV;
        }
        auto pick = [&](uint32_t j) {
            uint32_t v = INV;
#pragma unroll
            for (uint32_t t = 0; t < FM_SMALL + 2; ++t) v = (t == j) ? pr[t] : v;
            return v;
        };
        uint32_t a = 0, b = 1;
        bool done = false;
#pragma unroll
        for (uint32_t k = 0; k < FM_SMALL; ++k) {
            if (k < m && !done) {
                const uint32_t pa = pick(a);
                if (pa == INV) {
                    done = true;
                } else if (pa != sid[k]) {
                    fm_link(cs, mem[k], sid[k], g + a, pa, off);
                    a = b;
                    b = b + 1;
                } else {
                    const uint32_t pb = pick(b);
                    if (pb != INV) {
                        fm_link(cs, mem[k], sid[k], g + b, pb, off);
                        b = b + 1;
                    }
                }
            }
        }
        cs.head[g] = INV;
    }
#pragma unroll
    for (uint32_t k = 0; k < FM_SMALL; ++k)
        if (k < m) cs.grp[mem[k]] = INV;
}

// Groups of more than FM_SMALL members, up to FM_COOP_MAX, are sorted by
// the whole workgroup in LDS (bitonic) after its scan; the walk then reads
// the members' reads and the run's prefixes from LDS.  A group's members
// come from the group's row of the table (base FM_ROW, m), from the pool
// (base, m) or, for compacted lists of at most FM_GATHER_MAX suffixes (base
// INV), from a scan of the group marks, which costs less than following the
// member list one load at a time.
constexpr uint32_t FM_COOP_MAX = RESOLVE_COOP_MAX - 2;
struct FmCoop {
    uint64_t key[RESOLVE_COOP_MAX];  // sort keys; then the members' reads (u32) and the run's prefixes (u32)
    uint32_t mem[RESOLVE_COOP_MAX];
    uint32_t g[RESOLVE_COOP_LIST], base[RESOLVE_COOP_LIST], m[RESOLVE_COOP_LIST];
    uint32_t n, cnt;
};

template <class E>
__device__ __forceinline__ void fm_coop_groups(const ClaimState& cs, FmCoop& co, uint32_t off,
                                               const uint32_t* __restrict__ rows, uint32_t rw) {
    __syncthreads();
    const uint32_t nb = min(co.n, RESOLVE_COOP_LIST);
    const uint32_t ns = *cs.d_ns;
    uint32_t* sidl = reinterpret_cast<uint32_t*>(co.key);
    uint32_t* posl = sidl + RESOLVE_COOP_MAX;
    for (uint32_t gi = 0; gi < nb; ++gi) {
        const uint32_t g = co.g[gi];
        uint32_t m;
        if (co.base[gi] == INV) {
            if (threadIdx.x == 0) co.cnt = 0;
            __syncthreads();
            for (uint32_t x = threadIdx.x; x < ns; x += blockDim.x) {
                if (cs.grp[x] == g) {
                    const uint32_t q = atomicAdd(&co.cnt, 1u);
                    if (q < FM_COOP_MAX) co.mem[q] = x;
                }
            }
            __syncthreads();
            m = co.cnt;
            if (m > FM_COOP_MAX) {
                if (threadIdx.x == 0) fm_serial_group<E>(cs, g, off);
                __syncthreads();
                continue;
            }
        } else {
            m = co.m[gi];
            const uint32_t* arr = co.base[gi] == FM_ROW ? rows + (uint64_t)g * rw : cs.pool + co.base[gi];
            for (uint32_t x = threadIdx.x; x < m; x += blockDim.x) co.mem[x] = arr[x];
            __syncthreads();
        }
        uint32_t P = 1;
        while (P < m) P <<= 1;
        for (uint32_t x = threadIdx.x; x < P; x += blockDim.x) {
            if (x < m) {
                co.key[x] = rev_prefix_key<E>(cs.R, cs.W, cs.sl[co.mem[x]], off);
            } else {
                co.mem[x] = INV;  // padding sorts last
                co.key[x] = ~0ull;
            }
        }
        __syncthreads();
        for (uint32_t k = 2; k <= P; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t x = threadIdx.x; x < P; x += blockDim.x) {
                    const uint32_t y = x ^ j;
                    if (y > x) {
                        const uint32_t ux = co.mem[x], uy = co.mem[y];
                        const uint64_t kx = co.key[x], ky = co.key[y];
                        // the reads are loaded only to break a tie of the keys
                        const bool tie = ux != INV && uy != INV && kx == ky;
                        const uint32_t sx = tie ? cs.sl[ux] : INV, sy = tie ? cs.sl[uy] : INV;
                        const bool sw = ((x & k) == 0) ? fm_before<E>(cs, uy, ky, sy, ux, kx, sx, off)
                                                       : fm_before<E>(cs, ux, kx, sx, uy, ky, sy, off);
                        if (sw) {
                            co.mem[x] = uy;
                            co.mem[y] = ux;
                            co.key[x] = ky;
                            co.key[y] = kx;
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (cs.nf) {
            if (threadIdx.x == 0) claim_assign<E>(cs, co.mem, m, g, off);
        } else {
            // the members' reads, and the run's prefixes at positions 0 .. m + 1
            const uint32_t np = *cs.d_np, len = cs.L - off, n1 = min(E::BPW, len);
            const uint64_t* rs0 = cs.R + (uint64_t)cs.sl[co.mem[0]] * cs.W;
            const uint64_t sx00 = E::word(rs0, off, n1);
            __syncthreads();  // the keys are no longer read: their space holds sidl / posl
            for (uint32_t x = threadIdx.x; x < m + 2; x += blockDim.x) {
                if (x < m) sidl[x] = cs.sl[co.mem[x]];
                const uint32_t p = g + x;
                posl[x] =
                    (p < np && cmp_ps<E>(cs.R, cs.W, cs.pl[p], cs.pk[p], rs0, sx00, off, len) == 0) ? cs.pl[p] : INV;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                // the run is [g, e) of the sorted list: posl is INV from e - g on
                fm_walk_free(
                    cs, m, g, off, [&](uint32_t k) { return sidl[k]; }, [&](uint32_t k) { return co.mem[k]; },
                    [&](uint32_t j) { return posl[j]; });
                cs.head[g] = INV;
            }
        }
        __syncthreads();
        for (uint32_t x = threadIdx.x; x < m; x += blockDim.x) cs.grp[co.mem[x]] = INV;
        __syncthreads();
    }
}

// 64-bit value of lane (lane ^ j) / of lane k, as two 32-bit shuffles.
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, uint32_t j) {
    const uint32_t lo = __shfl_xor((uint32_t)v, (int)j, 64), hi = __shfl_xor((uint32_t)(v >> 32), (int)j, 64);
    return ((uint64_t)hi << 32) | lo;
}

// A group of up to FM_WAVE_MAX members whose members are all in its row,
// resolved by one wavefront (all 64 lanes call it with the same g, m): lane
// k loads member k from the row, the members are sorted by a bitonic
// network across the lanes (the order of fm_before; the reads are loaded
// only to break a tie of the keys), and the walk of fm_walk_free runs on
// every lane alike over the run's free prefixes (lane 0 writes the links).
// Compacted lists: every position of the run is free, lane j holds the
// j-th.  Lists in place (cs.nf): lane j checks position g + j (equal to the
// group's suffix and not linked), and the free ones are listed in lp (the
// wave's 128 words of LDS: positions, then prefixes) in order; if the run
// goes on past the 64 positions and they hold fewer than m + 2 free ones,
// nothing is done and false is returned (the caller resolves the group as
// before).
template <class E>
__device__ __forceinline__ bool fm_wave_group(const ClaimState& cs, uint32_t g, uint32_t m, uint32_t off,
                                              const uint32_t* __restrict__ row, uint32_t lane,
                                              uint32_t* __restrict__ lp) {
    uint32_t u = lane < m ? row[lane] : INV;
    uint32_t sid = u != INV ? cs.sl[u] : INV;
    uint64_t key = u != INV ? rev_prefix_key<E>(cs.R, cs.W, sid, off) : ~0ull;
#pragma unroll
    for (uint32_t k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const uint32_t pu = __shfl_xor(u, (int)j, 64), ps = __shfl_xor(sid, (int)j, 64);
            const uint64_t pkey = shfl_xor_u64(key, j);
            // the lower lane of a pair keeps the earlier member in an ascending run
            const bool first = ((lane & j) == 0) == ((lane & k) == 0);
            const bool sw = first ? fm_before<E>(cs, pu, pkey, ps, u, key, sid, off)
                                  : fm_before<E>(cs, u, key, sid, pu, pkey, ps, off);
            if (sw) {
                u = pu;
                sid = ps;
                key = pkey;
            }
        }
    }
    const uint32_t np = *cs.d_np, len = cs.L - off, n1 = min(E::BPW, len);
    const uint64_t* rs0 = cs.R + (uint64_t)__shfl(sid, 0, 64) * cs.W;
    const uint64_t sx00 = E::word(rs0, off, n1);
    const uint32_t p = g + lane;
    // pos(j) / px(j): the prefix and position of the run's j-th free
    // prefix, INV past the run; j <= m + 1 < 64
    uint32_t posl = INV;
    if (!cs.nf) {
        if (lane <= m + 1 && p < np && cmp_ps<E>(cs.R, cs.W, cs.pl[p], cs.pk[p], rs0, sx00, off, len) == 0)
            posl = cs.pl[p];
    } else {
        const bool eq = p < np && cmp_ps<E>(cs.R, cs.W, cs.pl[p], cs.pk[p], rs0, sx00, off, len) == 0;
        const bool fr = eq && cs.plinked[p] == 0;
        const uint64_t eb = __ballot(eq), fb = __ballot(fr);
        if (eb == ~0ull && (uint32_t)__popcll(fb) < m + 2) return false;
        const uint32_t nfree = (uint32_t)__popcll(fb);
        if (fr) {
            const uint32_t j = (uint32_t)__popcll(fb & lanemask_lt());
            lp[j] = p;
            lp[64 + j] = cs.pl[p];
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < nfree) posl = lp[64 + lane];
    }
    auto px = [&](uint32_t j) -> uint32_t { return cs.nf ? lp[j] : g + j; };
    // fm_walk_free; a, b <= m + 1 < 64
    uint32_t a = 0, b = 1;
    for (uint32_t k = 0; k < m; ++k) {
        const uint32_t pa = __shfl(posl, (int)a, 64);
        if (pa == INV) break;
        const uint32_t s = __shfl(sid, (int)k, 64), mk = __shfl(u, (int)k, 64);
        if (pa != s) {
            if (lane == 0) fm_link(cs, mk, s, px(a), pa, off);
            a = b;
            b = b + 1;
        } else {
            const uint32_t pb = __shfl(posl, (int)b, 64);
            if (pb != INV) {
                if (lane == 0) fm_link(cs, mk, s, px(b), pb, off);
                b = b + 1;
            }
        }
    }
    if (u != INV) cs.grp[u] = INV;
    if (lane == 0) cs.head[g] = INV;
    __builtin_amdgcn_wave_barrier();  // lp is read before the wave's next group overwrites it
    return true;
}

// Resolution of the groups k_fm_probe_s formed, found by their lead flags
// (cleared here).  A lone member links to its choice (claim_assign's walk
// gives the same prefix for one member); groups of up to FM_SMALL members
// are sorted in registers by the thread that found them, larger ones by
// the workgroup (up to FM_COOP_MAX and RESOLVE_COOP_LIST per workgroup;
// beyond, by that thread, as in claim_resolve_body).  Over compacted lists,
// groups whose members all sit in the row (up to FM_WAVE_MAX) are resolved
// by the wavefront that found them (fm_wave_group) instead.  Their members
// are walked in sorted order (claim_assign; fm_walk_free when the lists are
// compacted).  The list is split into one run of positions per wavefront,
// so that the groups of a short list spread over all of them.
template <class E>
__global__ void __launch_bounds__(256) k_fm_resolve_s(ClaimState cs, uint32_t off, uint8_t* __restrict__ lead,
                                                      uint32_t* __restrict__ gcnt, const uint32_t* __restrict__ gmem,
                                                      uint64_t gcap) {
    __shared__ FmCoop co;
    if (threadIdx.x == 0) co.n = 0;
    __syncthreads();
    const uint32_t ns = *cs.d_ns;
    const uint32_t rw = fm_row_width(gcap, *cs.d_np);
    // groups the wavefront resolves: over compacted lists 17 .. 62 members
    // (smaller ones by a lane in registers), over lists in place 2 .. 16
    // (a lane's walk would follow the skip pointers one load at a time)
    const uint32_t wmin = cs.nf ? 1u : FM_SMALL;
    const uint32_t wmax = cs.nf ? FM_SMALL : (gcap ? min(rw, FM_WAVE_MAX) : 0u);
    auto group = [&](uint32_t g, uint32_t m) {
        gcnt[g] = 0;
        if (m <= FM_SMALL) {
            uint32_t mem[FM_SMALL];
            const uint4* row = reinterpret_cast<const uint4*>(gmem + (uint64_t)g * rw);
#pragma unroll
            for (uint32_t q = 0; q < FM_SMALL / 4; ++q) {
                const uint4 v = row[q];
                mem[4 * q] = 4 * q < m ? v.x : INV;
                mem[4 * q + 1] = 4 * q + 1 < m ? v.y : INV;
                mem[4 * q + 2] = 4 * q + 2 < m ? v.z : INV;
                mem[4 * q + 3] = 4 * q + 3 < m ? v.w : INV;
            }
            const uint32_t u0 = mem[0];
            if (m == 1) {
                const uint32_t ch = cs.choice[u0];
                fm_link(cs, u0, cs.sl[u0], ch, cs.pl[ch], off);
                cs.grp[u0] = INV;
                cs.head[g] = INV;
            } else {
                fm_small_group<E>(cs, mem, m, g, off);
            }
            return;
        }
        const uint32_t u0 = cs.head[g];
        uint32_t u = u0;
        if (!cs.nf && gcap && m <= rw && m <= FM_COOP_MAX) {
            // every member is in the group's row
            const uint32_t slot = atomicAdd(&co.n, 1u);
            if (slot < RESOLVE_COOP_LIST) {
                co.g[slot] = g;
                co.base[slot] = FM_ROW;
                co.m[slot] = m;
                return;
            }
        }
        if (!cs.nf && ns <= FM_GATHER_MAX) {
            const uint32_t slot = atomicAdd(&co.n, 1u);
            if (slot < RESOLVE_COOP_LIST) {
                co.g[slot] = g;
                co.base[slot] = INV;
                return;
            }
        } else {
            if (m <= FM_COOP_MAX) {
                const uint32_t slot = atomicAdd(&co.n, 1u);
                if (slot < RESOLVE_COOP_LIST) {
                    const uint32_t base = atomicAdd(cs.pool_n, m);
                    uint32_t* arr = cs.pool + base;
                    m = 0;
                    for (u = u0; u != INV; u = cs.memb[u]) arr[m++] = u;
                    co.g[slot] = g;
                    co.base[slot] = base;
                    co.m[slot] = m;
                    return;
                }
            }
        }
        fm_serial_group<E>(cs, g, off);
    };
    // each wavefront scans its own run of positions
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wpb = blockDim.x >> 6;
    const uint64_t nwv = (uint64_t)gridDim.x * wpb;
    const uint64_t wid = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6);
    const uint64_t per = (ns + nwv - 1) / nwv;
    const uint64_t beg = wid * per;
    const uint64_t end = min<uint64_t>(beg + per, ns);
    // FM_SCAN positions at a time: their lead flags loaded together, the
    // leads listed in LDS (the cooperative sort's key space, unused until
    // fm_coop_groups), then resolved a lead per lane, so that the groups'
    // chains of loads overlap instead of following each 64 flags
    constexpr uint32_t FM_SCAN = 512;
    static_assert(4 * (FM_SCAN + 128) * 4 <= sizeof(co.key), "wave lead lists and run lists");
    uint32_t* wl = reinterpret_cast<uint32_t*>(co.key) + (threadIdx.x >> 6) * FM_SCAN;
    uint32_t* lp = reinterpret_cast<uint32_t*>(co.key) + 4 * FM_SCAN + (threadIdx.x >> 6) * 128;
    const uint64_t lt = lanemask_lt();
    for (uint64_t i0 = beg; i0 < end; i0 += FM_SCAN) {
        uint8_t f[FM_SCAN / 64];
#pragma unroll
        for (uint32_t r = 0; r < FM_SCAN / 64; ++r) {
            const uint64_t i = i0 + r * 64 + lane;
            f[r] = i < end ? lead[i] : (uint8_t)0;
        }
        uint32_t wn = 0;
#pragma unroll
        for (uint32_t r = 0; r < FM_SCAN / 64; ++r) {
            const bool hit = f[r] != 0;
            const uint64_t bal = __ballot(hit);
            if (hit) {
                const uint64_t i = i0 + r * 64 + lane;
                lead[i] = 0;
                wl[wn + (uint32_t)__popcll(bal & lt)] = (uint32_t)i;
            }
            wn += (uint32_t)__popcll(bal);
        }
        __builtin_amdgcn_wave_barrier();  // the list is read by other lanes of the wave
        for (uint32_t k0 = 0; k0 < wn; k0 += 64) {
            const uint32_t k = k0 + lane;
            uint32_t wg = INV, wm = 0;
            if (k < wn) {
                const uint32_t i = wl[k];
                const uint32_t g = cs.grp[i];
                const uint32_t m = gcnt[g];
                if (m > wmin && m <= wmax) {
                    wg = g;
                    wm = m;
                } else {
                    group(g, m);
                }
            }
            uint64_t bal = __ballot(wg != INV);
            while (bal) {
                const int kk = __builtin_ctzll(bal);
                bal &= bal - 1;
                const uint32_t gk = __shfl(wg, kk, 64), mk = __shfl(wm, kk, 64);
                const bool done = fm_wave_group<E>(cs, gk, mk, off, gmem + (uint64_t)gk * rw, lane, lp);
                if (done) {
                    if (lane == 0) gcnt[gk] = 0;
                } else if (lane == kk) {
                    group(gk, mk);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();  // before the next chunk's list overwrites this one
    }
    fm_coop_groups<E>(cs, co, off, gmem, rw);
}

// Minimizer-blocked filter of the prefixes' first 32 bases (the blocks and
// bits of k_prefix_hash), over a list whose length is a device counter.
__global__ void k_fm_prefix_filter(const uint64_t* __restrict__ R, const uint32_t* __restrict__ pl,
                                   const uint32_t* __restrict__ d_np, uint32_t W, uint32_t* __restrict__ mfilt,
                                   uint32_t bmask) {
    const uint32_t np = *d_np;
    GRID_STRIDE(i, np) {
        const uint64_t K = xw(R + (uint64_t)pl[i] * W, 0, 32);
        uint32_t* fb = mfilt + (uint64_t)ov_block(ov_minimizer(K, 32), bmask) * OVB_WORDS;
        const uint32_t f = ov_fbits(K), b1 = f & OVB_MASK, b2 = (f >> 10) & OVB_MASK;
        atomicOr(&fb[b1 >> 5], 1u << (b1 & 31u));
        atomicOr(&fb[b2 >> 5], 1u << (b2 & 31u));
    }
}

// ----------------------------------------------------------------- prune ---
// compress.cu:1539-1564
__global__ void k_prune_mark(const uint32_t* __restrict__ next, const uint32_t* __restrict__ prev,
                             const uint16_t* __restrict__ offset, uint64_t n, uint8_t* __restrict__ res) {
    GRID_STRIDE(r, n) {
        uint32_t nx = next[r], pv = prev[r];
        bool keep = (nx != INV && pv != INV) || (nx != INV && offset[r] == 0) || (pv != INV && offset[pv] == 0);
        res[r] = keep ? 1 : 0;
    }
}

// Links whose either end is dropped are cleared; the outcome does not depend
// on thread order (each cleared link is cleared from both ends).
__global__ void k_prune_unlink(uint32_t* __restrict__ next, uint32_t* __restrict__ prev,
                               const uint8_t* __restrict__ res, uint64_t n) {
    GRID_STRIDE(r, n) {
        if (res[r]) continue;
        uint32_t nx = next[r];
        if (nx != INV) {
            prev[nx] = INV;
            next[r] = INV;
        }
        uint32_t pv = prev[r];
        if (pv != INV) {
            next[pv] = INV;
            prev[r] = INV;
        }
    }
}

struct PredRes {
    const uint32_t* list;
    const uint8_t* res;
    bool want;
    __device__ bool operator()(uint32_t i) const { return (res[list[i]] != 0) == want; }
};

// ------------------------------------------------------- pseudogenome ----
// Writes nb bases of a read at PG position pos into a zeroed 2-bit array.
__global__ void k_pg_write(const uint64_t* __restrict__ R, uint32_t W, const uint32_t* __restrict__ ids,
                           const uint64_t* __restrict__ pos, const uint32_t* __restrict__ nbs, uint64_t n,
                           unsigned long long* __restrict__ pg) {
    GRID_STRIDE(i, n) {
        const uint64_t* rd = R + (uint64_t)ids[i] * W;
        uint64_t p = pos[i];
        uint32_t nb = nbs[i];
        for (uint32_t j = 0; j < nb; j += 32) {
            uint32_t c = min(32u, nb - j);
            uint64_t v = xw(rd, j, c);            // c bases, right-aligned
            uint64_t q = p + j;
            uint32_t o2 = (uint32_t)(q & 31u);     // bases already used in the first word
            // place the c bases at base offset o2 of word q/32
            uint64_t hi = (o2 + c <= 32) ? (v << (2 * (32 - o2 - c))) : (v >> (2 * (o2 + c - 32)));
            atomicOr(&pg[q >> 5], (unsigned long long)hi);
            if (o2 + c > 32) {
                uint32_t spill = o2 + c - 32;
                uint64_t lo = v << (2 * (32 - spill));
                atomicOr(&pg[(q >> 5) + 1], (unsigned long long)lo);
            }
        }
    }
}

// ------------------------------------------------------- N-read LQ -------
// 3-bit packing of N-reads for exact matching over A<C<G<N<T (Enc3): code
// A=0 C=1 G=2 N=3 T=4; other bytes fall outside the reference's alphabet
// (compress.cu:1205) and are coded as N.
__device__ __forceinline__ uint32_t enc5(uint8_t c) {
    return c == 'A' ? 0u : c == 'C' ? 1u : c == 'G' ? 2u : c == 'T' ? 4u : 3u;
}

__global__ void k_pack3(const uint8_t* __restrict__ ascii, uint64_t n, uint32_t L, uint32_t W3,
                        uint64_t* __restrict__ out) {
    GRID_STRIDE(t, n * W3) {
        uint64_t r = t / W3;
        uint32_t w = (uint32_t)(t - r * W3);
        uint32_t b0 = w * 21, nb = min(21u, L - b0);
        const uint8_t* s = ascii + r * L + b0;
        uint64_t v = 0;
        for (uint32_t j = 0; j < nb; ++j) v |= (uint64_t)enc5(s[j]) << (3 * (20 - j));
        out[t] = v;
    }
}

// Copies each record's contribution into an ASCII pseudogenome
// (compress.cu:167-217: the N-read pseudogenome is kept as text).
__global__ void k_pg_write_ascii(const uint8_t* __restrict__ ascii, uint32_t L, const uint32_t* __restrict__ ids,
                                 const uint64_t* __restrict__ pos, const uint32_t* __restrict__ nbs, uint64_t n,
                                 uint8_t* __restrict__ pg) {
    GRID_STRIDE(i, n) {
        const uint8_t* s = ascii + (uint64_t)ids[i] * L;
        uint8_t* d = pg + pos[i];
        for (uint32_t j = 0; j < nbs[i]; ++j) d[j] = s[j];
    }
}

// ---------------------------------------------------------------- mapping ---
constexpr uint32_t MAP_K = 28;       // Param.hpp:73 k1
constexpr uint32_t MAP_REF_STEP = 5; // ref_index_step1
constexpr uint32_t MAP_READ_STEP = 2;// read_index_step1
constexpr uint32_t MAP_TARGET = 1;   // target_mismatch_count1
constexpr uint32_t MAP_BUCKET = 12;  // bucket_limit

__global__ void k_map_hash(const uint64_t* __restrict__ pg, uint64_t start, uint64_t count, bool rc, uint32_t mask,
                           uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    GRID_STRIDE(i, count) {
        uint64_t km = xwu(pg, start + i * MAP_REF_STEP, MAP_K);
        if (rc) km = revcomp(km, MAP_K);
        keys[i] = (uint32_t)(murmur64(km) & mask);
        vals[i] = (uint32_t)i;
    }
}

// Bucket-inline index: bucket h is the 64-byte slot bk[16h .. 16h+15]
// holding its capped size in word 0 and its positions in words 1..12, so a
// probe reads one line instead of the CSR offsets' line plus the
// positions' line.  Used while H * 64 bytes stays within MAP_SLOT_MAX_H.
constexpr uint32_t MAP_SLOT = 16;
constexpr uint64_t MAP_SLOT_MAX_H = 1ull << 28;
// Both strands' slots of bucket h in one 128-byte line: forward at word 0,
// reverse complement at word MAP_SLOT (k_map_both).
constexpr uint32_t MAP_SLOT2 = 2 * MAP_SLOT;

// One thread per bucket writes its slot: the capped size and the first
// min(n,12) occurrences (ascending i, as sorted) in descending order, the
// order of k_map_fill.  Words past the bucket's positions are not written:
// readers never look past the size.  16-byte stores, the first always, the
// others only when the bucket reaches them.
__global__ void k_map_slots(const uint32_t* __restrict__ vals, const uint32_t* __restrict__ krf, uint64_t H,
                            uint32_t* __restrict__ bk) {
    // four consecutive lanes write the four 16-byte pieces of one slot: the
    // head always, the others as far as the size reaches
    GRID_STRIDE(t, H * 4) {
        const uint64_t h = t >> 2;
        const uint32_t q = (uint32_t)t & 3u;
        const uint32_t lo = krf[h];
        const uint32_t c = min(krf[h + 1] - lo, MAP_BUCKET);
        if (q == 0 || c > 4 * q - 1) {
            // slot word 0: the size; word 1 + p: position p, in descending order
            uint32_t w[4];
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                const uint32_t word = 4 * q + k;
                w[k] = word == 0 ? c : (word - 1 < c ? vals[lo + c - word] : 0u);
            }
            *reinterpret_cast<uint4*>(bk + h * MAP_SLOT + 4 * q) = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
}

// Both halves of the two-strand table (MAP_SLOT2) in one pass: bucket h's
// forward slot from (vals0, krf0) and reverse-complement slot from (vals1,
// krf1), laid out as k_map_slots lays them out.  Eight consecutive lanes
// write the eight 16-byte pieces of one bucket's line, all of them (zeros
// past the sizes), so a wave writes eight whole consecutive lines; partial
// lines cost more at HBM than the zeros do.
__global__ void k_map_slots2(const uint32_t* __restrict__ vals0, const uint32_t* __restrict__ krf0,
                             const uint32_t* __restrict__ vals1, const uint32_t* __restrict__ krf1, uint64_t H,
                             uint32_t* __restrict__ bk) {
    GRID_STRIDE(t, H * 8) {
        const uint64_t h = t >> 3;
        const uint32_t sd = (uint32_t)(t >> 2) & 1u, q = (uint32_t)t & 3u;  // strand, piece of its slot
        const uint32_t* vals = sd ? vals1 : vals0;
        const uint32_t* krf = sd ? krf1 : krf0;
        const uint32_t lo = krf[h];
        const uint32_t c = min(krf[h + 1] - lo, MAP_BUCKET);
        {
            // slot word 0: the size; word 1 + p: position p, in descending order
            uint32_t w[4];
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                const uint32_t word = 4 * q + k;
                w[k] = word == 0 ? c : (word - 1 < c ? vals[lo + c - word] : 0u);
            }
            *reinterpret_cast<uint4*>(bk + h * MAP_SLOT2 + sd * MAP_SLOT + 4 * q) = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
}

// Compact two-strand index: one 32-byte slot per bucket, so that the table
// of a 2^26-bucket index is 2 GiB instead of 8 GiB of 128-byte lines (random
// probes into a table of 4 GiB or more miss the address-translation caches
// and run at about two thirds of the rate of probes into one of 1-2 GiB on
// MI355X).  Slot word 0: the forward size (bits 0-3) and the
// reverse-complement size (bits 4-7), each capped at MAP_BUCKET.  A bucket
// with at most CS_INLINE positions in all keeps them in words 1..7, the
// forward ones first, each strand in its bucket order (as k_map_slots); a
// larger one (0.2% of the probes that reach an indexed position on a 100 Mbp
// genome at 30x) keeps in word 1 the offset of its positions, in the same
// order, in the overflow array.
constexpr uint32_t CS_INLINE = 7;

// Overflow sizes: the bucket's positions when they do not fit its slot,
// else 0; entries [0, H] (entry H = 0, so that an exclusive scan ends with
// the total).
__global__ void k_cslot_count(const uint32_t* __restrict__ krf0, const uint32_t* __restrict__ krf1, uint64_t H,
                              uint32_t* __restrict__ ovc) {
    GRID_STRIDE(h, H + 1) {
        uint32_t t = 0;
        if (h < H) {
            t = min(krf0[h + 1] - krf0[h], MAP_BUCKET) + min(krf1[h + 1] - krf1[h], MAP_BUCKET);
            if (t <= CS_INLINE) t = 0;
        }
        ovc[h] = t;
    }
}

// One thread per bucket writes its slot, tab[2h] and tab[2h+1] (and its
// overflow positions at ovo[h], the exclusive scan of k_cslot_count).
__global__ void k_cslot_write(const uint32_t* __restrict__ vals0, const uint32_t* __restrict__ krf0,
                              const uint32_t* __restrict__ vals1, const uint32_t* __restrict__ krf1, uint64_t H,
                              const uint32_t* __restrict__ ovo, uint4* __restrict__ tab, uint32_t* __restrict__ ovf) {
    GRID_STRIDE(h, H) {
        const uint32_t l0 = krf0[h], c0 = min(krf0[h + 1] - l0, MAP_BUCKET);
        const uint32_t l1 = krf1[h], c1 = min(krf1[h + 1] - l1, MAP_BUCKET);
        const uint32_t t = c0 + c1;
        // position k of the bucket: forward first; the first min(n,12)
        // occurrences of each strand in descending order
        auto at = [&](uint32_t k) -> uint32_t {
            return k < c0 ? vals0[l0 + c0 - 1 - k] : vals1[l1 + c1 - 1 - (k - c0)];
        };
        uint32_t w[8] = {c0 | (c1 << 4), 0u, 0u, 0u, 0u, 0u, 0u, 0u};
        if (t <= CS_INLINE) {
#pragma unroll
            for (uint32_t k = 0; k < CS_INLINE; ++k)
                if (k < t) w[1 + k] = at(k);
        } else {
            const uint32_t o = ovo[h];
            w[1] = o;
            for (uint32_t k = 0; k < t; ++k) ovf[o + k] = at(k);
        }
        tab[2 * h] = make_uint4(w[0], w[1], w[2], w[3]);
        tab[2 * h + 1] = make_uint4(w[4], w[5], w[6], w[7]);
    }
}

// Capped bucket sizes: min(run length, 12), entries [0, H+2).
__global__ void k_map_caps(const uint32_t* __restrict__ krf, uint64_t H, uint32_t* __restrict__ caps) {
    GRID_STRIDE(b, H + 2) {
        uint32_t c = 0;
        if (b < H) c = min(krf[b + 1] - krf[b], MAP_BUCKET);
        caps[b] = c;
    }
}

// The first min(n,12) occurrences (ascending i, as sorted) land in
// descending order (the sequential meaning of compress.cu:369-406).
__global__ void k_map_fill(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint64_t count,
                           const uint32_t* __restrict__ krf, const uint32_t* __restrict__ kr,
                           uint32_t* __restrict__ pos) {
    GRID_STRIDE(t, count) {
        uint32_t h = keys[t];
        uint32_t q = (uint32_t)t - krf[h];
        uint32_t c = min(krf[h + 1] - krf[h], MAP_BUCKET);
        if (q < c) pos[kr[h] + c - 1 - q] = vals[t];
    }
}

__device__ __forceinline__ uint32_t fwd_mm(const uint64_t* __restrict__ ref, uint64_t o1,
                                           const uint64_t* __restrict__ rd, uint32_t L, uint32_t bound) {
    uint32_t mm = 0, i = 0;
    for (i = 0; (i + 32 < L) && (mm <= bound); i += 32) mm += mm_words(xwu(ref, o1 + i, 32), xw(rd, i, 32));
    if ((i < L) && (mm <= bound)) mm += mm_words(xwu(ref, o1 + i, L - i), xw(rd, i, L - i));
    return mm;
}

__device__ __forceinline__ uint32_t rc_mm(const uint64_t* __restrict__ ref, uint64_t o1,
                                          const uint64_t* __restrict__ rd, uint32_t L, uint32_t bound) {
    uint32_t mm = 0, i = 0;
    for (i = 0; (i + 32 < L) && (mm <= bound); i += 32)
        mm += mm_words(xwu(ref, o1 + i, 32), revcomp(xw(rd, L - (i + 32), 32), 32));
    if ((i < L) && (mm <= bound)) mm += mm_words(xwu(ref, o1 + i, L - i), revcomp(xw(rd, 0, L - i), L - i));
    return mm;
}

__device__ __forceinline__ uint64_t kmer_ascii(const uint8_t* s, uint32_t k) {
    uint64_t bits = 0;
    for (uint32_t i = 0; i < k; ++i) bits |= (uint64_t)(enc_base(s[i]) & 3u) << (2 * (k - 1 - i));
    return bits;
}

__device__ __forceinline__ uint8_t base_char(uint32_t c) {
    return (uint8_t)(c == 0 ? 'A' : c == 1 ? 'C' : c == 2 ? 'G' : 'T');
}

struct MapArgs {
    const uint64_t* pg;
    uint64_t pg_len;
    const uint32_t* kr;
    const uint32_t* pos;
    uint32_t mask;
    uint64_t start;
    const uint64_t* R;         // packed reads, by global id
    uint32_t W, L;
    const uint32_t* un_ids;    // unmatched non-N ids
    uint64_t n_un;
    const uint8_t* nascii;     // N-reads ASCII
    uint64_t n_N;
    uint8_t* last;
    uint64_t* word;            // record word per mapping slot
    // [0] active reads, [1] k-mer probes, [2] verified candidates,
    // [3] reads settled (<= 1 mismatch) at their first probe, [4] probes and
    // [5] candidates of reads left above 1 mismatch by this pass
    unsigned long long* work;
    uint32_t max_mm = 0;       // floor(L / 5) (used by the probe-range kernels)
    const uint32_t* bk = nullptr;  // bucket-inline index (kr / pos unused when set)
    // slot h of the strand's index starts at word h * bk_stride + bk_off of
    // bk: one strand per table (MAP_SLOT, 0), or both strands in one table
    // (MAP_SLOT2, MAP_SLOT * strand)
    uint32_t bk_stride = 16;
    uint32_t bk_off = 0;
    // compact two-strand index (k_cslot_write): bk holds one 16-byte slot
    // per bucket, ovf the positions of buckets that do not fit their slot
    uint32_t cs = 0;
    const uint32_t* ovf = nullptr;
    // global ids of the N-reads (their packed rows in R) and the masks of
    // their bases outside A, C, G, T (k_nmask): with them, N-reads are
    // compared word by word (nread_count, k_map_both's N-read instance)
    const uint32_t* nids = nullptr;
    const uint64_t* nmask = nullptr;
};

// Masks of the N-reads' bases outside A, C, G, T (upper case: the
// characters that mismatch every pseudogenome base in the reference's
// character compare, compress.cu:411-521), in the packed layout (01 in the
// base's field): W words per N-read, in N-read order.
__global__ void k_nmask(const uint8_t* __restrict__ nascii, uint64_t nN, uint32_t L, uint32_t W,
                        uint64_t* __restrict__ out) {
    GRID_STRIDE(t, nN * W) {
        const uint64_t r = t / W;
        const uint32_t k = (uint32_t)(t - r * W);
        const uint8_t* ra = nascii + r * L + 32 * k;
        const uint32_t n = min(32u, L - 32 * k);
        uint64_t m = 0;
        for (uint32_t b = 0; b < n; ++b) {
            const uint8_t ch = ra[b];
            if (ch != 'A' && ch != 'C' && ch != 'G' && ch != 'T') m |= 1ull << (62u - 2u * b);
        }
        out[t] = m;
    }
}

// Mismatches of two 2-bit words, with the bases marked in m (01 in the
// base's field) counted whatever their codes.
__device__ __forceinline__ uint32_t mm_words_m(uint64_t a, uint64_t b, uint64_t m) {
    const uint64_t x = a ^ b;
    return (uint32_t)__popcll(((x & 0xAAAAAAAAAAAAAAAAULL) >> 1) | (x & 0x5555555555555555ULL) | m);
}

// Both strands' slots of a bucket of a two-strand index: the 128-byte line
// form (MAP_SLOT2: f and r are the slots' first 16 bytes, sl the line) or
// the compact form (CS: f and r are the two halves of the 32-byte slot, sl
// the bucket's overflow positions when it has any).
struct Bk2 {
    uint4 f, r;
    const uint32_t* sl;
    uint32_t nf, nr;  // sizes (nr 0 when the reverse slot was not asked for)
    uint32_t nt;      // compact form: both sizes together
};
// Position p of the forward (rc false) or reverse-complement slot.
template <bool CS>
__device__ __forceinline__ uint32_t bk2_pos(const Bk2& b, bool rc, uint32_t p) {
    if (CS) {
        const uint32_t k = rc ? b.nf + p : p;
        if (b.nt > CS_INLINE) return ld32(b.sl + k);
        switch (k) {
            case 0: return b.f.y;
            case 1: return b.f.z;
            case 2: return b.f.w;
            case 3: return b.r.x;
            case 4: return b.r.y;
            case 5: return b.r.z;
            default: return b.r.w;
        }
    }
    const uint4& hd = rc ? b.r : b.f;
    return p == 0 ? hd.y : p == 1 ? hd.z : p == 2 ? hd.w : ld32(b.sl + (rc ? MAP_SLOT : 0u) + 1 + p);
}

// Bucket h of the mapping index: its capped size and its positions.
struct MapBucket {
    uint32_t n;
    const uint32_t* p;
};
__device__ __forceinline__ MapBucket map_bucket(const MapArgs& a, uint32_t h) {
    if (a.bk) {
        const uint32_t* s = a.bk + (uint64_t)h * a.bk_stride + a.bk_off;
        return MapBucket{ld32(s), s + 1};
    }
    const uint32_t b = a.kr[h], e = a.kr[h + 1];
    return MapBucket{min(e - b, MAP_BUCKET), a.pos + b};
}

// Bucket h of a two-strand index; the reverse-complement slot only when
// want_r (its size reads 0 otherwise).
template <bool CS>
__device__ __forceinline__ Bk2 bk2_fetch(const MapArgs& a, uint32_t h, bool want_r) {
    Bk2 b;
    b.r = make_uint4(0, 0, 0, 0);
    if (CS) {
        const uint4* sl = reinterpret_cast<const uint4*>(a.bk) + 2 * (uint64_t)h;
        b.f = sl[0];
        b.r = sl[1];
        b.nf = b.f.x & 15u;
        const uint32_t nr = (b.f.x >> 4) & 15u;
        b.nt = b.nf + nr;
        b.nr = want_r ? nr : 0u;
        b.sl = b.nt > CS_INLINE ? a.ovf + b.f.y : a.ovf;
        return b;
    }
    b.sl = a.bk + (uint64_t)h * MAP_SLOT2;
    b.f = *reinterpret_cast<const uint4*>(b.sl);
    if (want_r) b.r = *reinterpret_cast<const uint4*>(b.sl + MAP_SLOT);
    b.nf = b.f.x;
    b.nr = b.r.x;
    b.nt = 0;
    return b;
}

__device__ __forceinline__ void wave_add_u64(unsigned long long* dst, uint32_t v) {
    uint64_t s = v;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if (__lane_id() == 0 && s) atomicAdd(dst, (unsigned long long)s);
}

// Positions verified earlier in the same strand pass.  Skipping a repeat is
// exact: its mismatch count is the same, and since `last` only decreases a
// candidate rejected (or accepted) once cannot be strictly better later.
struct SeenCache {
    uint64_t s0 = ~0ull, s1 = ~0ull, s2 = ~0ull, s3 = ~0ull;
    __device__ __forceinline__ bool check_insert(uint64_t mp) {
        if (mp == s0 || mp == s1 || mp == s2 || mp == s3) return true;
        s3 = s2;
        s2 = s1;
        s1 = s0;
        s0 = mp;
        return false;
    }
};

// Mismatches of an N-read (packed row rd, mask row mk) against the window at
// mp, 32 bases at a time, the masked bases counted whatever their codes:
// the count of the reference's character loop (compress.cu:437-450 with
// util.cuh:271-281), which stops at `last`, so counts of last or more come
// back as last.
template <bool RC>
__device__ __forceinline__ uint32_t nread_count(const MapArgs& a, const uint64_t* __restrict__ rd,
                                                const uint64_t* __restrict__ mk, uint64_t mp, uint32_t last) {
    const uint32_t L = a.L;
    uint32_t c = 0;
    for (uint32_t o = 0; o < L && c < last; o += 32) {
        const uint32_t n = min(32u, L - o);
        const uint64_t win = xwu(a.pg, mp + o, n);
        uint64_t q, m;
        if (RC) {
            q = revcomp(xw(rd, L - o - n, n), n);
            m = Enc2::reversed(xw(mk, L - o - n, n), n);
        } else {
            q = xw(rd, o, n);
            m = xw(mk, o, n);
        }
        c += mm_words_m(win, q, m);
    }
    return min(c, last);
}

// One N-read through the k_map loop: with the masks (a.nmask) word by word
// (nread_count), else base by base from its characters.
template <bool RC>
__device__ __forceinline__ void map_nread(const MapArgs& a, uint64_t i, uint32_t& last, uint64_t& w,
                                          SeenCache& seen, uint32_t& n_probe, uint32_t& n_cand) {
    const uint8_t* ra = a.nascii + (i - a.n_un) * a.L;
    const uint64_t* rd = a.nmask ? a.R + (uint64_t)a.nids[i - a.n_un] * a.W : nullptr;
    const uint64_t* mk = a.nmask ? a.nmask + (i - a.n_un) * a.W : nullptr;
    for (uint32_t j = 0; j + MAP_K <= a.L; j += MAP_READ_STEP) {
        ++n_probe;
        // the packed row holds the characters as enc_base & 3, as kmer_ascii
        uint32_t h = (uint32_t)(murmur64(rd ? xw(rd, j, MAP_K) : kmer_ascii(ra + j, MAP_K)) & a.mask);
        MapBucket B{0u, nullptr};
        Bk2 cb{};
        if (a.cs) {
            cb = bk2_fetch<true>(a, h, RC);
            B.n = RC ? cb.nr : cb.nf;
        } else {
            B = map_bucket(a, h);
        }
        if (B.n == 0) continue;
        const uint64_t ls = RC ? (uint64_t)(a.L - MAP_K - j) : (uint64_t)j;
        for (uint32_t p = 0; p < B.n; ++p) {
            const uint32_t pv = a.cs ? bk2_pos<true>(cb, RC, p) : ld32(B.p + p);
            uint64_t rp = a.start + (uint64_t)pv * MAP_REF_STEP;
            if (rp < ls) continue;
            uint64_t mp = rp - ls;
            if (mp + a.L > a.pg_len) continue;
            if (seen.check_insert(mp)) continue;
            ++n_cand;
            uint32_t c = 0;
            if (rd) {
                c = nread_count<RC>(a, rd, mk, mp, last);
            } else {
                for (uint32_t o = 0; o < a.L; ++o) {
                    uint8_t rb = base_char(base2(a.pg, mp + o));
                    uint8_t qb = RC ? rev_base_ascii(ra[a.L - 1 - o]) : ra[o];
                    if (rb != qb) {
                        c++;
                        if (c >= last) break;
                    }
                }
            }
            if (c < last) {
                w = 1ull | (RC ? 2ull : 0ull) | (mp << 2);
                last = c;
                if (last <= MAP_TARGET) return;
            }
        }
    }
}

// compress.cu:411-521: one thread per read; candidates in bucket order,
// a strictly better mismatch count replaces the record, stop at <= 1.
template <bool RC>
__global__ void __launch_bounds__(256) k_map(MapArgs a) {
    uint32_t n_act = 0, n_probe = 0, n_cand = 0, n_first = 0, n_open_probe = 0, n_open_cand = 0;
    GRID_STRIDE(i, a.n_un + a.n_N) {
        uint32_t last = a.last[i];
        if (last <= MAP_TARGET) continue;
        ++n_act;
        const uint32_t probe0 = n_probe, cand0 = n_cand;
        const bool isN = i >= a.n_un;
        uint64_t w = a.word[i];
        SeenCache seen;
        if (!isN) {
            const uint64_t* rd = a.R + (uint64_t)a.un_ids[i] * a.W;
            for (uint32_t j = 0; j + MAP_K <= a.L; j += MAP_READ_STEP) {
                ++n_probe;
                uint32_t h = (uint32_t)(murmur64(xw(rd, j, MAP_K)) & a.mask);
                const MapBucket B = map_bucket(a, h);
                if (B.n == 0) continue;
                const uint64_t ls = RC ? (uint64_t)(a.L - MAP_K - j) : (uint64_t)j;
                for (uint32_t p = 0; p < B.n; ++p) {
                    uint64_t rp = a.start + (uint64_t)ld32(B.p + p) * MAP_REF_STEP;
                    if (rp < ls) continue;
                    uint64_t mp = rp - ls;
                    if (mp + a.L > a.pg_len) continue;
                    if (seen.check_insert(mp)) continue;
                    ++n_cand;
                    uint32_t c = RC ? rc_mm(a.pg, mp, rd, a.L, last - 1) : fwd_mm(a.pg, mp, rd, a.L, last - 1);
                    if (c < last) {
                        w = (RC ? 2ull : 0ull) | (mp << 2);
                        last = c;
                        if (last <= MAP_TARGET) goto done;
                    }
                }
            }
        } else {
            map_nread<RC>(a, i, last, w, seen, n_probe, n_cand);
        }
    done:
        a.last[i] = (uint8_t)last;
        a.word[i] = w;
        if (last <= MAP_TARGET) {
            n_first += (n_probe - probe0 == 1) ? 1u : 0u;
        } else {
            n_open_probe += n_probe - probe0;
            n_open_cand += n_cand - cand0;
        }
    }
    wave_add_u64(&a.work[0], n_act);
    wave_add_u64(&a.work[1], n_probe);
    wave_add_u64(&a.work[2], n_cand);
    wave_add_u64(&a.work[3], n_first);
    wave_add_u64(&a.work[4], n_open_probe);
    wave_add_u64(&a.work[5], n_open_cand);
}

// ---- k_map with the read in registers (bucket-inline index only) ----------
// The same loop, candidate order, seen-cache and early exits as k_map, so
// the same records and counters; what changes is how often a thread waits
// on memory.  The read's words (and, on the RC strand, the chunks of its
// reverse complement) are loaded once into registers; the bucket heads of
// MAP_GROUP probes are requested together; a candidate's pseudogenome
// window is requested in one go (W + 1 independent 8-byte loads) instead
// of chunk by chunk.  WM is the register capacity in words (W <= WM).
constexpr uint32_t MAP_GROUP = 4;

// a[i] for a register array and a run-time i (0 if i >= WM), by selects.
template <uint32_t WM, bool K32>
__global__ void __launch_bounds__(256) k_ov_verdicts(const uint64_t* __restrict__ R, const uint32_t* __restrict__ sl,
                                                     uint32_t ns, uint32_t W, uint32_t k, uint32_t kip,
                                                     uint32_t max_off, const uint32_t* __restrict__ mfilt,
                                                     uint32_t bmask, uint64_t* __restrict__ vb, uint64_t nw,
                                                     const uint32_t* __restrict__ d_ns) {
    if (d_ns) ns = *d_ns;  // the list length as a device counter (ns is then its bound)
    __shared__ uint32_t blocks[256 * (OVB_WORDS + 1)];  // odd stride: spreads banks
    uint32_t* mine = blocks + threadIdx.x * (OVB_WORDS + 1);
    uint32_t held = 0xFFFFFFFFu;  // block in the LDS slot
    auto verdict = [&](uint64_t K, uint32_t mv, bool valid, uint64_t i, uint32_t off) {
        const uint32_t blk = ov_block(mv, bmask);
        if (blk != held) {
            const uint4* src = reinterpret_cast<const uint4*>(mfilt + (uint64_t)blk * OVB_WORDS);
#pragma unroll
            for (uint32_t q = 0; q < OVB_WORDS / 4; ++q) {
                const uint4 v = src[q];
                mine[4 * q] = v.x;
                mine[4 * q + 1] = v.y;
                mine[4 * q + 2] = v.z;
                mine[4 * q + 3] = v.w;
            }
            held = blk;
        }
        const uint32_t f = ov_fbits(K), b1 = f & OVB_MASK, b2 = (f >> 10) & OVB_MASK;
        const bool pass = valid && ((mine[b1 >> 5] >> (b1 & 31u)) & (mine[b2 >> 5] >> (b2 & 31u)) & 1u) != 0;
        const uint64_t bal = __ballot(pass);
        // a wave past the list's end (lane 0 beyond it) has no word to write
        if (__lane_id() == 0 && valid) vb[(uint64_t)(off - 1) * nw + (i >> 6)] = bal;
    };
    for (uint64_t base = (uint64_t)blockIdx.x * 256; base < ns; base += (uint64_t)gridDim.x * 256) {
        const uint64_t i = base + threadIdx.x;
        const bool valid = i < ns;
        uint64_t r[WM];
        const uint64_t* rd = R + (uint64_t)(valid ? sl[i] : 0u) * W;
#pragma unroll
        for (uint32_t w = 0; w < WM; ++w) r[w] = (valid && w < W) ? rd[w] : 0ull;
        if constexpr (K32) {
            constexpr uint32_t NM = 32 - OVM + 1;
            uint32_t hc[NM];  // m-mer hashes of the current chunk (positions x0 .. x0 + NM - 1)
            for (uint32_t o0 = 1; o0 < max_off; o0 += NM) {
                const uint32_t x0 = kip + o0;
                const uint64_t A = xw_pick(r, x0, 32), B = xw_pick(r, x0 + 32, 32);
                // the 32 bases from base t of A:B (t <= 33)
                auto win = [&](uint32_t t) -> uint64_t {
                    if (t == 0) return A;
                    if (t < 32) return (A << (2 * t)) | (B >> (64 - 2 * t));
                    return B << (2 * (t - 32));
                };
                if (o0 == 1) {
#pragma unroll
                    for (uint32_t t = 0; t < NM; ++t) hc[t] = ov_mhash((uint32_t)(win(t) >> 32));
                }
                uint32_t hn[NM];  // the next chunk's
#pragma unroll
                for (uint32_t t = 0; t < NM; ++t) hn[t] = ov_mhash((uint32_t)(win(NM + t) >> 32));
#pragma unroll
                for (uint32_t t = NM - 1; t-- > 0;) hc[t] = min(hc[t], hc[t + 1]);  // suffix minima
                uint32_t pm = 0xFFFFFFFFu;  // prefix minimum of hn[0 .. j)
#pragma unroll
                for (uint32_t j = 0; j < NM; ++j) {
                    const uint32_t off = o0 + j;
                    if (off >= max_off) break;
                    verdict(win(j), min(hc[j], pm), valid, i, off);
                    pm = min(pm, hn[j]);
                }
#pragma unroll
                for (uint32_t t = 0; t < NM; ++t) hc[t] = hn[t];
            }
        } else {
            for (uint32_t off = 1; off < max_off; ++off) {
                const uint64_t K = xw_pick(r, kip + off, k);
                verdict(K, ov_minimizer(K, k), valid, i, off);
            }
        }
    }
}

template <bool RC, uint32_t WM>
__global__ void __launch_bounds__(256) k_map_fast(MapArgs a) {
    uint32_t n_act = 0, n_probe = 0, n_cand = 0, n_first = 0, n_open_probe = 0, n_open_cand = 0;
    const uint32_t L = a.L, W = a.W;
    const uint32_t np = (L - MAP_K) / MAP_READ_STEP + 1;
    GRID_STRIDE(i, a.n_un + a.n_N) {
        uint32_t last = a.last[i];
        if (last <= MAP_TARGET) continue;
        ++n_act;
        const uint32_t probe0 = n_probe, cand0 = n_cand;
        uint64_t w = a.word[i];
        SeenCache seen;
        if (i < a.n_un) {
            const uint64_t* rd = a.R + (uint64_t)a.un_ids[i] * W;
            uint64_t r[WM], q[WM];
#pragma unroll
            for (uint32_t k = 0; k < WM; ++k) r[k] = k < W ? ld64(rd + k) : 0ull;
            // query chunk k: bases [32k, 32k + n_k) of the read or of its
            // reverse complement, right-aligned (the chunking of fwd_mm / rc_mm)
#pragma unroll
            for (uint32_t k = 0; k < WM; ++k) {
                q[k] = 0;
                if (32 * k < L) {
                    const uint32_t n = min(32u, L - 32 * k);
                    q[k] = RC ? revcomp(xw_pick(r, L - 32 * k - n, n), n) : xw_pick(r, 32 * k, n);
                }
            }
            for (uint32_t t0 = 0; t0 < np; t0 += MAP_GROUP) {
                uint4 hd[MAP_GROUP];
                uint32_t hh[MAP_GROUP];
#pragma unroll
                for (uint32_t u = 0; u < MAP_GROUP; ++u) {
                    hd[u] = make_uint4(0, 0, 0, 0);
                    hh[u] = 0;
                    if (t0 + u < np) {
                        hh[u] = (uint32_t)(murmur64(xw_pick(r, MAP_READ_STEP * (t0 + u), MAP_K)) & a.mask);
                        hd[u] = *reinterpret_cast<const uint4*>(a.bk + (uint64_t)hh[u] * MAP_SLOT);
                    }
                }
#pragma unroll
                for (uint32_t u = 0; u < MAP_GROUP; ++u) {
                    const uint32_t t = t0 + u;
                    if (t >= np) break;
                    ++n_probe;
                    const uint32_t nb = hd[u].x;
                    const uint32_t j = MAP_READ_STEP * t;
                    const uint64_t ls = RC ? (uint64_t)(L - MAP_K - j) : (uint64_t)j;
                    for (uint32_t p = 0; p < nb; ++p) {
                        const uint32_t pv = p == 0   ? hd[u].y
                                            : p == 1 ? hd[u].z
                                            : p == 2 ? hd[u].w
                                                     : ld32(a.bk + (uint64_t)hh[u] * MAP_SLOT + 1 + p);
                        const uint64_t rp = a.start + (uint64_t)pv * MAP_REF_STEP;
                        if (rp < ls) continue;
                        const uint64_t mp = rp - ls;
                        if (mp + L > a.pg_len) continue;
                        if (seen.check_insert(mp)) continue;
                        ++n_cand;
                        // the window's W + 1 words, requested together
                        const uint64_t* g0 = a.pg + (mp >> 5);
                        uint64_t g[WM + 1];
#pragma unroll
                        for (uint32_t k = 0; k <= WM; ++k) g[k] = k <= W ? ld64(g0 + k) : 0ull;
                        const uint32_t sh = (uint32_t)(mp & 31u) << 1;
                        uint32_t c = 0;
                        const uint32_t bound = last - 1;
#pragma unroll
                        for (uint32_t k = 0; k < WM; ++k) {
                            if (32 * k < L && c <= bound) {
                                const uint32_t n = min(32u, L - 32 * k);
                                const uint64_t win = sh ? (g[k] << sh) | (g[k + 1] >> (64u - sh)) : g[k];
                                c += mm_words(win >> ((32u - n) << 1), q[k]);
                            }
                        }
                        if (c < last) {
                            w = (RC ? 2ull : 0ull) | (mp << 2);
                            last = c;
                            if (last <= MAP_TARGET) goto done;
                        }
                    }
                }
            }
        } else {
            map_nread<RC>(a, i, last, w, seen, n_probe, n_cand);
        }
    done:
        a.last[i] = (uint8_t)last;
        a.word[i] = w;
        if (last <= MAP_TARGET) {
            n_first += (n_probe - probe0 == 1) ? 1u : 0u;
        } else {
            n_open_probe += n_probe - probe0;
            n_open_cand += n_cand - cand0;
        }
    }
    wave_add_u64(&a.work[0], n_act);
    wave_add_u64(&a.work[1], n_probe);
    wave_add_u64(&a.work[2], n_cand);
    wave_add_u64(&a.work[3], n_first);
    wave_add_u64(&a.work[4], n_open_probe);
    wave_add_u64(&a.work[5], n_open_cand);
}

// SeenCache over 32-bit positions (pseudogenomes shorter than 2^32 bases).
struct SeenCache32 {
    uint32_t s0 = ~0u, s1 = ~0u, s2 = ~0u, s3 = ~0u;
    __device__ __forceinline__ bool check_insert(uint64_t mp64) {
        const uint32_t mp = (uint32_t)mp64;
        if (mp == s0 || mp == s1 || mp == s2 || mp == s3) return true;
        s3 = s2;
        s2 = s1;
        s1 = s0;
        s0 = mp;
        return false;
    }
};

// Both strand passes of a read in one loop over its probes, on the
// two-strand table (MAP_SLOT2): each probe fetches one line that holds the
// bucket's forward and reverse-complement slots, where the separate passes
// fetch two.  The result and the work counters are those of the forward
// pass followed by the reverse-complement pass (k_map_fast<false> then
// <true>):
//  - the forward state is exactly the forward pass's; when it reaches
//    MAP_TARGET the read is done and the reverse pass never counts;
//  - the reverse pass runs alongside with its own best Rb, accepting c <
//    min(Rb, F) where F is the forward best so far.  Since F only falls and
//    the sequential pass starts from the final F, a candidate with c >= F
//    can never be the reverse winner; the reverse winner is the first
//    candidate reaching the least such c, as in the sequential pass, and it
//    replaces the forward result when it is below the final F.  Its partial
//    sums stop at min(Rb, F) - 1 >= 1, so counts at or below MAP_TARGET are
//    exact and it stops at the same candidate as the sequential pass;
//  - each pass keeps its own SeenCache, and the reverse work is added to
//    the counters only when the forward pass ends above MAP_TARGET.
//
// NR: the instance for the N-reads (indexes [n_un, n_un + n_N)), launched
// after the one for the other reads.  The reference compares an N-read
// character by character (compress.cu:411-521 with util.cuh:271-281): a
// read character that is not one of A, C, G, T (upper case) mismatches
// every pseudogenome base, and k-mers hash the read packed with such
// characters as code enc_base & 3 (N as A).  Here the read is that packed
// row (k_pack_w packs it so), compared word by word like the other reads,
// and a mask (k_nmask) marks its bases outside A, C, G, T, which count as
// mismatches whatever their codes (in the reverse-complement pass the
// complement of such a character is again outside A, C, G, T, and the mask
// is reversed with the read).  So an N-read's counts, results and work counters are
// those of the character loop, at the cost of the other reads.
// Early end of a strand pass that cannot change its result.  The pass keeps
// the first window of least count, and replaces it only with a window of
// fewer mismatches.  Let B be the count to beat: a window q with at most
// B - 1 mismatches matches the read exactly in every k-mer window
// [j, j + MAP_K) that holds none of them.  Probe j (j = 2t) can list q only
// if q + j (forward; q + L - MAP_K - j on the reverse strand) is a multiple
// of MAP_REF_STEP, which fixes t mod 5: q's probes are those of one class,
// 10 bases apart, so one mismatch spoils at most ceil(MAP_K / 10) = 3 of
// them.  A probe whose bucket holds fewer than MAP_BUCKET positions lists
// every indexed position of its k-mer's hash.  So once every class has had
// more than 3 (B - 1) such complete probes, every q that could beat B has
// already been a candidate (and was counted then, against a bound at least
// as large): the remaining probes cannot change the pass's result.  clean
// holds the complete probes per class, 4 bits each.  Only the work counts
// change (fewer probes and candidates than the exhaustive loop).
constexpr uint32_t MAP_SETTLE_KILL = (MAP_K + 2 * MAP_REF_STEP - 1) / (2 * MAP_REF_STEP);  // 3
constexpr uint32_t MAP_SETTLE_MIN = 5 * (MAP_SETTLE_KILL + 1);  // probes before the first class can settle
static_assert(MAP_READ_STEP == 2 && MAP_REF_STEP == 5, "probe classes assume steps 2 and 5");
__device__ __forceinline__ bool map_settled(uint32_t clean, uint32_t B) {
    if (B <= 1) return true;
    const uint32_t need = MAP_SETTLE_KILL * (B - 1) + 1;
    if (need > 15) return false;
    uint32_t m = clean & 15u;
#pragma unroll
    for (uint32_t c = 1; c < 5; ++c) m = min(m, (clean >> (4 * c)) & 15u);
    return m >= need;
}

template <uint32_t WM, bool CS, bool NR>
__global__ void __launch_bounds__(256, 7) k_map_both(MapArgs a) {
    constexpr uint32_t G = 1;  // probes whose slot heads are requested together
    uint32_t n_act = 0, n_probe = 0, n_cand = 0, n_first = 0, n_open_probe = 0, n_open_cand = 0;
    const uint32_t L = a.L, W = a.W;
    const uint32_t np = (L - MAP_K) / MAP_READ_STEP + 1;
    const uint64_t i1 = NR ? a.n_un + a.n_N : a.n_un;
    const uint64_t flagN = NR ? 1ull : 0ull;  // is_contain_N bit of the record word
    for (uint64_t i = (NR ? a.n_un : 0) + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < i1;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t last = a.last[i];
        if (last <= MAP_TARGET) continue;
        uint64_t w = a.word[i];
        uint64_t qf[WM], qr[WM];
        uint64_t mf[NR ? WM : 1], mr[NR ? WM : 1];  // N-reads: bases outside A, C, G, T
        {
            const uint64_t* rd = a.R + (uint64_t)(NR ? a.nids[i - a.n_un] : a.un_ids[i]) * W;
            uint64_t r[WM];
#pragma unroll
            for (uint32_t k = 0; k < WM; ++k) r[k] = k < W ? ld64(rd + k) : 0ull;
#pragma unroll
            for (uint32_t k = 0; k < WM; ++k) {
                qf[k] = qr[k] = 0;
                if (32 * k < L) {
                    const uint32_t n = min(32u, L - 32 * k);
                    qf[k] = xw_pick(r, 32 * k, n);
                    qr[k] = revcomp(xw_pick(r, L - 32 * k - n, n), n);
                }
            }
            mf[0] = mr[0] = 0;
            if constexpr (NR) {
                // the read's mask (k_nmask)
                uint64_t mk[WM];
                const uint64_t* mrow = a.nmask + (i - a.n_un) * W;
#pragma unroll
                for (uint32_t k = 0; k < WM; ++k) mk[k] = k < W ? ld64(mrow + k) : 0ull;
#pragma unroll
                for (uint32_t k = 0; k < WM; ++k) {
                    mf[k] = mr[k] = 0;
                    if (32 * k < L) {
                        const uint32_t n = min(32u, L - 32 * k);
                        mf[k] = xw_pick(mk, 32 * k, n);
                        // reversed, not complemented
                        mr[k] = Enc2::reversed(xw_pick(mk, L - 32 * k - n, n), n);
                    }
                }
            }
        }
        // the read's words again from qf: all full except the last chunk,
        // which qf holds right-aligned
        const uint32_t nlast = L - 32 * (W - 1);
        auto rword = [&](uint32_t k) -> uint64_t {
            const uint64_t v = pick(qf, k);
            return (k == W - 1 && nlast < 32) ? v << (2 * (32 - nlast)) : v;
        };
        auto kmer = [&](uint32_t off) -> uint64_t {  // MAP_K bases from base off (as xw_pick)
            const uint32_t o1 = off >> 5;
            const uint32_t o2 = (off & 31u) << 1;
            const uint64_t left = rword(o1) << o2;
            if ((64u - o2) >= (MAP_K << 1)) return left >> ((32u - MAP_K) << 1);
            const uint64_t right = (rword(o1 + 1) >> (62u - o2)) >> 2;
            return (left | right) >> ((32u - MAP_K) << 1);
        };
        // mismatches of query q against the window at mp, stopping once above
        // bound.  The first 32 bases come from two words (one 128-byte line
        // in 15 of 16 cases); the rest of the window is fetched only when
        // they stay within the bound, which hash-bucket neighbours rarely do.
        auto count = [&](const uint64_t (&q)[WM], const uint64_t (&m)[NR ? WM : 1], uint64_t mp,
                         uint32_t bound) -> uint32_t {
            const uint64_t* g0 = a.pg + (mp >> 5);
            uint64_t g[WM + 1];
            g[0] = ld64(g0);
            g[1] = ld64(g0 + 1);  // W >= 1
            const uint32_t sh = (uint32_t)(mp & 31u) << 1;
            uint32_t c;
            {
                const uint32_t n = min(32u, L);
                const uint64_t win = sh ? (g[0] << sh) | (g[1] >> (64u - sh)) : g[0];
                if constexpr (NR)
                    c = mm_words_m(win >> ((32u - n) << 1), q[0], m[0]);
                else
                    c = mm_words(win >> ((32u - n) << 1), q[0]);
            }
            if (c > bound || L <= 32) return c;
#pragma unroll
            for (uint32_t k = 2; k <= WM; ++k) g[k] = k <= W ? ld64(g0 + k) : 0ull;
#pragma unroll
            for (uint32_t k = 1; k < WM; ++k) {
                if (32 * k < L && c <= bound) {
                    const uint32_t n = min(32u, L - 32 * k);
                    const uint64_t win = sh ? (g[k] << sh) | (g[k + 1] >> (64u - sh)) : g[k];
                    if constexpr (NR)
                        c += mm_words_m(win >> ((32u - n) << 1), q[k], m[NR ? k : 0]);
                    else
                        c += mm_words(win >> ((32u - n) << 1), q[k]);
                }
            }
            return c;
        };
        uint32_t F = last, Rb = last;  // forward best, reverse best (below F when it counts)
        uint64_t wF = w, wR = w;
        bool doneR = false, tgtR = false;  // reverse pass over; over because it reached MAP_TARGET
        bool doneF = false;                // forward pass over above MAP_TARGET (map_settled)
        uint32_t cleanF = 0, cleanR = 0;   // complete buckets probed per probe class (map_settled)
        uint32_t cls = 0;                  // class of probe t: t mod 5
        uint32_t pF = 0, cF = 0, pR = 0, cR = 0;  // work of each pass
        SeenCache32 seenF, seenR;  // positions < pg_len < 2^32
        for (uint32_t t0 = 0; t0 < np; t0 += G) {
            uint4 hf[G], hr[G];
            uint32_t hh[G];
            Bk2 bb[G];  // compact index (CS)
#pragma unroll
            for (uint32_t u = 0; u < G; ++u) {
                hf[u] = hr[u] = make_uint4(0, 0, 0, 0);
                hh[u] = 0;
                if (t0 + u < np) {
                    hh[u] = (uint32_t)(murmur64(kmer(MAP_READ_STEP * (t0 + u))) & a.mask);
                    if constexpr (CS) {
                        bb[u] = bk2_fetch<true>(a, hh[u], !doneR);
                    } else {
                        const uint32_t* sl = a.bk + (uint64_t)hh[u] * MAP_SLOT2;
                        hf[u] = *reinterpret_cast<const uint4*>(sl);
                        if (!doneR) hr[u] = *reinterpret_cast<const uint4*>(sl + MAP_SLOT);
                    }
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < G; ++u) {
                const uint32_t t = t0 + u;
                if (t >= np) break;
                const uint32_t j = MAP_READ_STEP * t;
                const uint32_t* sl = a.bk + (uint64_t)hh[u] * MAP_SLOT2;
                // forward pass, probe t
                if (!doneF) {
                    ++pF;
                    uint32_t nb;
                    if constexpr (CS)
                        nb = bb[u].nf;
                    else
                        nb = hf[u].x;
                    if (nb < MAP_BUCKET) cleanF += 1u << (4 * cls);
                    const uint64_t ls = (uint64_t)j;
                    for (uint32_t p = 0; p < nb; ++p) {
                        uint32_t pv;
                        if constexpr (CS)
                            pv = bk2_pos<true>(bb[u], false, p);
                        else
                            pv = p == 0 ? hf[u].y : p == 1 ? hf[u].z : p == 2 ? hf[u].w : ld32(sl + 1 + p);
                        const uint64_t rp = a.start + (uint64_t)pv * MAP_REF_STEP;
                        if (rp < ls) continue;
                        const uint64_t mp = rp - ls;
                        if (mp + L > a.pg_len) continue;
                        if (seenF.check_insert(mp)) continue;
                        ++cF;
                        const uint32_t c = count(qf, mf, mp, F - 1);
                        if (c < F) {
                            wF = flagN | (mp << 2);
                            F = c;
                            if (F <= MAP_TARGET) goto fwd_done;
                        }
                    }
                }
                // reverse pass, probe t
                if (!doneR) {
                    ++pR;
                    uint32_t nb;
                    if constexpr (CS)
                        nb = bb[u].nr;
                    else
                        nb = hr[u].x;
                    if (nb < MAP_BUCKET) cleanR += 1u << (4 * cls);
                    const uint64_t ls = (uint64_t)(L - MAP_K - j);
                    for (uint32_t p = 0; p < nb; ++p) {
                        uint32_t pv;
                        if constexpr (CS)
                            pv = bk2_pos<true>(bb[u], true, p);
                        else
                            pv = p == 0   ? hr[u].y
                                 : p == 1 ? hr[u].z
                                 : p == 2 ? hr[u].w
                                          : ld32(sl + MAP_SLOT + 1 + p);
                        const uint64_t rp = a.start + (uint64_t)pv * MAP_REF_STEP;
                        if (rp < ls) continue;
                        const uint64_t mp = rp - ls;
                        if (mp + L > a.pg_len) continue;
                        if (seenR.check_insert(mp)) continue;
                        ++cR;
                        const uint32_t lim = min(Rb, F);
                        const uint32_t c = count(qr, mr, mp, lim - 1);
                        if (c < lim) {
                            wR = flagN | 2ull | (mp << 2);
                            Rb = c;
                            if (Rb <= MAP_TARGET) {
                                doneR = tgtR = true;
                                break;
                            }
                        }
                    }
                }
                cls = cls == 4 ? 0u : cls + 1u;
                // a pass whose answer no window it has not seen can change
                // ends here (map_settled).  Forward: a window must beat F,
                // and reach the reverse best Rb (a tie goes to the forward
                // pass) unless it reaches MAP_TARGET, which ends the read
                // before the reverse pass counts.  Reverse: it must beat
                // min(Rb, F).
                if (t + 1 >= MAP_SETTLE_MIN) {
                    if (!doneF && map_settled(cleanF, min(F, max(Rb + 1, MAP_TARGET + 1)))) doneF = true;
                    if (!doneR && map_settled(cleanR, min(Rb, F))) doneR = true;
                }
            }
            if (doneF && doneR) break;
        }
        // forward pass ended above MAP_TARGET: the reverse pass counts
        n_act += 2;
        n_probe += pF + pR;
        n_cand += cF + cR;
        n_open_probe += pF;
        n_open_cand += cF;
        if (tgtR) {
            n_first += (pR == 1) ? 1u : 0u;
        } else {
            n_open_probe += pR;
            n_open_cand += cR;
        }
        if (Rb < F) {
            last = Rb;
            w = wR;
        } else {
            last = F;
            w = wF;
        }
        a.last[i] = (uint8_t)last;
        a.word[i] = w;
        continue;
    fwd_done:
        // forward pass reached MAP_TARGET: its result; no reverse pass
        ++n_act;
        n_probe += pF;
        n_cand += cF;
        n_first += (pF == 1) ? 1u : 0u;
        a.last[i] = (uint8_t)F;
        a.word[i] = wF;
    }
    wave_add_u64(&a.work[0], n_act);
    wave_add_u64(&a.work[1], n_probe);
    wave_add_u64(&a.work[2], n_cand);
    wave_add_u64(&a.work[3], n_first);
    wave_add_u64(&a.work[4], n_open_probe);
    wave_add_u64(&a.work[5], n_open_cand);
}

// Alignment records in the 16-byte layout of compress.cu:68-77, built on the
// device so that one copy delivers them: read_id = id_map[rid[i]] (or
// rid[i]), word = (v[i] << shift) | flag.
__global__ void k_rec_aos(const uint32_t* __restrict__ rid, const uint64_t* __restrict__ v,
                          const uint32_t* __restrict__ id_map, uint64_t n, uint32_t shift, uint64_t flag,
                          curc_alignment* __restrict__ out) {
    GRID_STRIDE(i, n) {
        curc_alignment r;
        r.read_id = id_map ? id_map[rid[i]] : rid[i];
        r.pad = 0;
        r.word = (v[i] << shift) | flag;
        out[i] = r;
    }
}

__global__ void k_map_init(uint8_t* __restrict__ last, uint64_t* __restrict__ word, uint64_t n_un, uint64_t n,
                           uint32_t max_mm) {
    GRID_STRIDE(i, n) {
        last[i] = (uint8_t)(max_mm + 1);
        word[i] = (i >= n_un) ? 1ull : 0ull;
    }
}

struct PredMapped {
    const uint8_t* last;
    uint32_t max_mm;
    uint32_t base;
    bool want;
    __device__ bool operator()(uint32_t i) const { return (last[base + i] <= max_mm) == want; }
};
struct EmitMapRec {
    const uint8_t* last;
    const uint64_t* word;
    const uint32_t* un_ids;
    uint32_t n_un;
    uint32_t* rid;
    uint64_t* rword;
    uint8_t* mis;
    __device__ void operator()(uint32_t i, uint32_t j) const {
        rid[j] = (i < n_un) ? un_ids[i] : i - n_un;
        rword[j] = word[i];
        mis[j] = last[i];
    }
};
struct EmitUnmapped {
    const uint32_t* un_ids;  // null for N-reads (emit the index)
    uint32_t* out;
    __device__ void operator()(uint32_t i, uint32_t j) const { out[j] = un_ids ? un_ids[i] : i; }
};

}  // namespace curc
