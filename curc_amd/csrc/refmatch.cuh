// Pseudogenome self-match (RefMatcher, compress.cu:597-1001) on gfx950.
//
// The source is the HQ pseudogenome (2-bit); a destination (the HQ text
// itself, the LQ text, or the N-read LQ text in ASCII) is reverse
// complemented, probed every third position with a k=36 maRushPrime1 hash
// against an index of every fifth source position, and verified seeds are
// extended word-wise (32 bases per compare) in both directions.
//
// The reference's racy parts are fixed as follows: buckets hold the first 12
// source positions (ascending), visited in descending order; one result slot
// per segment (binary) or position (character) keeps results in collection
// order, truncated at the capacity there.  The sequential emission pass of
// compress.cu:955-993 is computed in parallel: with matches sorted by
// destination start, the accepted set is the chain head -> nxt -> nxt ...
// where nxt(i) is the first later match whose end - 50 reaches the end of i;
// the chain is marked by pointer doubling.
#pragma once
#include "kernels.cuh"

namespace curc {

constexpr uint32_t RM_K = 36;           // compress.cu:678 kmer_size
constexpr uint32_t RM_SRC_STEP = 5;     // compress.cu:676 src_step
constexpr uint32_t RM_DEST_STEP = 3;    // compress.cu:677 dest_step
constexpr uint32_t RM_TARGET = 50;      // compress.cu:679 target_match_len
constexpr uint32_t RM_SEG = 4;          // compress.cu:750 dest_segement_step
constexpr uint32_t RM_TILE = 256;       // tile of the "next reaching match" search

// ASCII of a 2-bit code: A=65 C=67 G=71 T=84.
__device__ __forceinline__ uint32_t rm_char(uint32_t c) { return 65u + ((0x13060200u >> (8u * c)) & 0xFFu); }

// helper.cpp:253-268 (complementsLut): IUPAC complements, lower case to upper,
// anything else to 0.
__device__ __forceinline__ uint8_t rm_comp(uint8_t c) {
    switch (c) {
        case 'A': case 'a': case 'U': case 'u': return 'T';
        case 'C': case 'c': return 'G';
        case 'G': case 'g': return 'C';
        case 'T': case 't': return 'A';
        case 'N': case 'n': return 'N';
        case 'Y': case 'y': return 'R';
        case 'R': case 'r': return 'Y';
        case 'K': case 'k': return 'M';
        case 'M': case 'm': return 'K';
        case 'B': case 'b': return 'V';
        case 'D': case 'd': return 'H';
        case 'H': case 'h': return 'D';
        case 'V': case 'v': return 'B';
        default: return 0;
    }
}

// maRushPrime1 (compress.cu:597-615) of 36 bases given as the first 32
// (one right-aligned word) and the last 4 (8 bits): the hash reads 4 ASCII
// characters at a time as a little-endian u32.
__device__ __forceinline__ uint32_t rm_hash2(uint64_t hi, uint32_t lo4) {
    uint64_t h = RM_K;
#pragma unroll
    for (uint32_t j = 0; j < RM_K / 4; ++j) {
        const uint32_t g = j < 8 ? (uint32_t)(hi >> (56u - 8u * j)) & 0xFFu : lo4;
        uint32_t k = rm_char((g >> 6) & 3u) | rm_char((g >> 4) & 3u) << 8 | rm_char((g >> 2) & 3u) << 16 |
                     rm_char(g & 3u) << 24;
        k += j;
        h ^= k;
        h *= 171717ull;
    }
    return (uint32_t)h;
}

// compress.cu:617-629 over 36 ASCII bytes.
__device__ __forceinline__ uint32_t rm_hash_ascii(const uint8_t* __restrict__ s) {
    uint64_t h = RM_K;
#pragma unroll
    for (uint32_t j = 0; j < RM_K / 4; ++j) {
        uint32_t k = (uint32_t)ld8(s + 4 * j) | (uint32_t)ld8(s + 4 * j + 1) << 8 |
                     (uint32_t)ld8(s + 4 * j + 2) << 16 | (uint32_t)ld8(s + 4 * j + 3) << 24;
        k += j;
        h ^= k;
        h *= 171717ull;
    }
    return (uint32_t)h;
}

// Index keys of source positions 5i (compress.cu:712-741).
__global__ void k_rm_hash(const uint64_t* __restrict__ src, uint64_t count, uint32_t mask,
                          uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    GRID_STRIDE(i, count) {
        const uint64_t p = i * RM_SRC_STEP;
        keys[i] = rm_hash2(xwu(src, p, 32), (uint32_t)xwu(src, p + 32, 4)) & mask;
        vals[i] = (uint32_t)i;
    }
}

// Reverse complement of an n-base 2-bit text into `out` (nwords =
// ceil(n/32)+1 words; bases past n are 0).
__global__ void k_rm_rc2(const uint64_t* __restrict__ in, uint64_t n, uint64_t* __restrict__ out, uint64_t nwords) {
    GRID_STRIDE(w, nwords) {
        uint64_t v = 0;
        const uint64_t first = w * 32;
        if (first < n) {
            const uint32_t nb = (uint32_t)min<uint64_t>(32, n - first);
            v = revcomp(xwu(in, n - first - nb, nb), nb) << (64u - 2u * nb);
        }
        out[w] = v;
    }
}

__global__ void k_rm_rc_ascii(const uint8_t* __restrict__ in, uint64_t n, uint8_t* __restrict__ out) {
    GRID_STRIDE(i, n) out[i] = rm_comp(in[n - 1 - i]);
}

// 2-bit packing of an n-base ASCII text (one word per thread; bases other
// than A/C/G/T pack as 0, as in encode, compress.cu:2151-2180).
__global__ void k_rm_pack_text(const uint8_t* __restrict__ text, uint64_t n, uint64_t* __restrict__ out,
                               uint64_t nwords) {
    GRID_STRIDE(w, nwords) {
        uint64_t v = 0;
        for (uint32_t k = 0; k < 32; ++k) {
            const uint64_t i = w * 32 + k;
            const uint32_t c = i < n ? (enc_base(text[i]) & 3u) : 0u;
            v |= (uint64_t)c << (62u - 2u * k);
        }
        out[w] = v;
    }
}

struct RmArgs {
    const uint64_t* src;   // HQ pseudogenome, 2-bit
    uint64_t src_len;
    const uint32_t* kr;    // CSR bucket ranges [H+2]
    const uint32_t* pos;   // source indices (x5 = position), descending per bucket
    uint32_t mask;
    const uint64_t* rc2;   // reverse-complemented destination, 2-bit (binary path)
    const uint8_t* rca;    // reverse-complemented destination, ASCII (character path)
    uint64_t dest_len;
    uint64_t n_slots;      // segments (binary) or positions (character)
    bool is_ref;
    uint64_t* r_src;
    uint64_t* r_len;
    uint64_t* r_dest;
    uint8_t* flag;
    unsigned long long* work;  // [0] candidates verified, [1] bases compared in extensions
};

// Longest common run of `lim` bases going right from (a, b) over 2-bit texts.
__device__ __forceinline__ uint64_t rm_lcp2(const uint64_t* __restrict__ x, uint64_t a, const uint64_t* __restrict__ y,
                                            uint64_t b, uint64_t lim) {
    uint64_t r = 0;
    while (r < lim) {
        const uint32_t nb = (uint32_t)min<uint64_t>(32, lim - r);
        const uint64_t d = xwu(x, a + r, nb) ^ xwu(y, b + r, nb);
        if (d) return r + (uint64_t)((__clzll(d) - (64 - 2 * (int)nb)) >> 1);
        r += nb;
    }
    return r;
}

// Longest common run going left from (a, b) exclusive, at most `lim` bases.
__device__ __forceinline__ uint64_t rm_lcs2(const uint64_t* __restrict__ x, uint64_t a, const uint64_t* __restrict__ y,
                                            uint64_t b, uint64_t lim) {
    uint64_t l = 0;
    while (l < lim) {
        const uint32_t nb = (uint32_t)min<uint64_t>(32, lim - l);
        const uint64_t d = xwu(x, a - l - nb, nb) ^ xwu(y, b - l - nb, nb);
        if (d) return l + (uint64_t)(__builtin_ctzll(d) >> 1);
        l += nb;
    }
    return l;
}

// Seed at destination position dp against bucket candidates (compress.cu:
// 755-803): the first candidate whose 36-mer equals the seed and whose
// extension spans at least 50 bases.  The left walk of the reference stops
// at a text start without counting that base, hence `bnd - 1` there.
__device__ bool rm_probe2(const RmArgs& a, uint64_t dp, uint64_t& o_src, uint64_t& o_len, uint64_t& o_dest,
                          uint32_t& n_cand, uint64_t& n_ext) {
    const uint64_t khi = xwu(a.rc2, dp, 32);
    const uint32_t klo = (uint32_t)xwu(a.rc2, dp + 32, 4);
    const uint32_t h = rm_hash2(khi, klo) & a.mask;
    const uint32_t qb = a.kr[h], qe = a.kr[h + 1];
    for (uint32_t q = qb; q < qe; ++q) {
        const uint64_t sp = (uint64_t)a.pos[q] * RM_SRC_STEP;
        if (a.is_ref && a.dest_len - sp < dp) continue;  // compress.cu:764
        if (xwu(a.src, sp, 32) != khi || (uint32_t)xwu(a.src, sp + 32, 4) != klo) continue;
        ++n_cand;
        const uint64_t lim = min(a.dest_len - dp - RM_K, a.src_len - sp - RM_K);
        const uint64_t r = rm_lcp2(a.rc2, dp + RM_K, a.src, sp + RM_K, lim);
        const uint64_t bnd = min(dp, sp);
        const uint64_t l = rm_lcs2(a.rc2, dp, a.src, sp, bnd);
        n_ext += r + l;
        const uint64_t start = l < bnd ? dp - l : dp - bnd + 1;
        const uint64_t right = dp + RM_K + r;
        if (right - start >= RM_TARGET) {
            o_src = sp + start - dp;
            o_len = right - start;
            o_dest = start;
            return true;
        }
    }
    return false;
}

// Binary destination: one thread per segment of 4 probe positions; the
// first verified match ends the segment (compress.cu:746-807).
__global__ void __launch_bounds__(256) k_rm_match2(RmArgs a) {
    uint32_t n_cand = 0;
    uint64_t n_ext = 0;
    GRID_STRIDE(g, a.n_slots) {
        bool hit = false;
        uint64_t s = 0, l = 0, d = 0;
        for (uint32_t k = 0; k < RM_SEG && !hit; ++k) {
            const uint64_t dp = (g * RM_SEG + k) * RM_DEST_STEP;
            if (dp + RM_K > a.dest_len) break;  // the 36-mer would run past the text
            hit = rm_probe2(a, dp, s, l, d, n_cand, n_ext);
        }
        a.flag[g] = hit ? 1 : 0;
        if (hit) {
            a.r_src[g] = s;
            a.r_len[g] = l;
            a.r_dest[g] = d;
        }
    }
    if (a.work) {
        wave_add_u64(a.work, n_cand);
        wave_add_u64(a.work + 1, (uint32_t)min<uint64_t>(n_ext, 0xFFFFFFFFull));
    }
}

// Character destination (N-read LQ text): one thread per probe position,
// byte compares against the decoded source (compress.cu:809-853).
__global__ void __launch_bounds__(256) k_rm_match_ascii(RmArgs a) {
    uint32_t n_cand = 0;
    uint64_t n_ext = 0;
    GRID_STRIDE(i, a.n_slots) {
        const uint64_t dp = i * RM_DEST_STEP;
        const uint8_t* s = a.rca + dp;
        const uint32_t h = rm_hash_ascii(s) & a.mask;
        const uint32_t qb = a.kr[h], qe = a.kr[h + 1];
        bool hit = false;
        for (uint32_t q = qb; q < qe && !hit; ++q) {
            const uint64_t sp = (uint64_t)a.pos[q] * RM_SRC_STEP;
            bool eq = true;
            for (uint32_t j = 0; j < RM_K && eq; ++j) eq = ld8(s + j) == rm_char(base2(a.src, sp + j));
            if (!eq) continue;
            ++n_cand;
            const uint64_t lim = min(a.dest_len - dp - RM_K, a.src_len - sp - RM_K);
            uint64_t r = 0;
            while (r < lim && ld8(a.rca + dp + RM_K + r) == rm_char(base2(a.src, sp + RM_K + r))) ++r;
            const uint64_t bnd = min(dp, sp);
            uint64_t l = 0;
            while (l < bnd && ld8(a.rca + dp - 1 - l) == rm_char(base2(a.src, sp - 1 - l))) ++l;
            n_ext += r + l;
            const uint64_t start = l < bnd ? dp - l : dp - bnd + 1;
            const uint64_t right = dp + RM_K + r;
            if (right - start >= RM_TARGET) {
                a.r_src[i] = sp + start - dp;
                a.r_len[i] = right - start;
                a.r_dest[i] = start;
                hit = true;
            }
        }
        a.flag[i] = hit ? 1 : 0;
    }
    if (a.work) {
        wave_add_u64(a.work, n_cand);
        wave_add_u64(a.work + 1, (uint32_t)min<uint64_t>(n_ext, 0xFFFFFFFFull));
    }
}

struct PredFlag8 {
    const uint8_t* f;
    __device__ bool operator()(uint32_t i) const { return f[i] != 0; }
};

struct EmitTriple {
    const uint64_t *s, *l, *d;
    uint64_t *os, *ol, *od;
    __device__ void operator()(uint32_t i, uint32_t j) const {
        os[j] = s[i];
        ol[j] = l[i];
        od[j] = d[i];
    }
};

// compress.cu:914-930: forward destination coordinates; for the self match
// the earlier copy becomes the source and an overlap is split in half.
__global__ void k_rm_transform(uint64_t* __restrict__ s, uint64_t* __restrict__ l, uint64_t* __restrict__ d,
                               uint64_t n, uint64_t dest_len, bool is_ref, uint64_t* __restrict__ key,
                               uint32_t* __restrict__ idx) {
    GRID_STRIDE(i, n) {
        uint64_t ps = s[i], ln = l[i], pd = dest_len - (d[i] + l[i]);
        if (is_ref) {
            if (ps > pd) {
                const uint64_t t = ps;
                ps = pd;
                pd = t;
            }
            const uint64_t end_src = ps + ln;
            if (end_src > pd) {
                const uint64_t margin = (end_src - pd + 1) / 2;
                ln -= margin;
                pd += margin;
            }
        }
        s[i] = ps;
        l[i] = ln;
        d[i] = pd;
        key[i] = ln;
        idx[i] = (uint32_t)i;
    }
}

__global__ void k_rm_gather_key(const uint64_t* __restrict__ v, const uint32_t* __restrict__ idx, uint64_t n,
                                uint64_t* __restrict__ key) {
    GRID_STRIDE(i, n) key[i] = v[idx[i]];
}

// Sorted triple i survives unique() when it differs from triple i-1.
struct PredUniqueTriple {
    const uint64_t *s, *l, *d;
    const uint32_t* idx;
    __device__ bool operator()(uint32_t i) const {
        if (i == 0) return true;
        const uint32_t a = idx[i], b = idx[i - 1];
        return s[a] != s[b] || l[a] != l[b] || d[a] != d[b];
    }
};

struct EmitTripleIdx {
    const uint64_t *s, *l, *d;
    const uint32_t* idx;
    uint64_t *os, *ol, *od;
    __device__ void operator()(uint32_t i, uint32_t j) const {
        const uint32_t a = idx[i];
        os[j] = s[a];
        ol[j] = l[a];
        od[j] = d[a];
    }
};

// Reach threshold T = end - 49 for matches of at least 50 bases (0 = never
// accepted): a match is accepted after a previous end P iff P < T.  Tile
// maxima of T feed the next-match search.
__global__ void __launch_bounds__(RM_TILE) k_rm_reach(const uint64_t* __restrict__ l, const uint64_t* __restrict__ d,
                                                      uint64_t n, uint64_t* __restrict__ T,
                                                      uint64_t* __restrict__ tmax) {
    __shared__ uint64_t red[RM_TILE];
    const uint64_t i = (uint64_t)blockIdx.x * RM_TILE + threadIdx.x;
    uint64_t t = 0;
    if (i < n && l[i] >= RM_TARGET) t = d[i] + l[i] - (RM_TARGET - 1);
    if (i < n) T[i] = t;
    red[threadIdx.x] = t;
    __syncthreads();
    for (uint32_t s = RM_TILE / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) tmax[blockIdx.x] = red[0];
}

// nxt[i] for i in [0, n]: node n is the start (end 0); the answer is the
// first j > i (j >= 0 for the start) with T[j] > end(i), or n + 1.
__global__ void k_rm_next(const uint64_t* __restrict__ l, const uint64_t* __restrict__ d,
                          const uint64_t* __restrict__ T, const uint64_t* __restrict__ tmax, uint64_t n,
                          uint32_t* __restrict__ nxt) {
    GRID_STRIDE(i, n + 2) {
        if (i == n + 1) {
            nxt[i] = (uint32_t)(n + 1);
            continue;
        }
        const uint64_t e = i < n ? d[i] + l[i] : 0;
        uint64_t j = i < n ? i + 1 : 0;
        const uint64_t ntile = (n + RM_TILE - 1) / RM_TILE;
        uint64_t res = n + 1;
        // rest of j's tile
        const uint64_t tend = min(n, (j / RM_TILE + 1) * RM_TILE);
        for (; j < tend; ++j)
            if (T[j] > e) {
                res = j;
                break;
            }
        if (res > n) {
            uint64_t b = (j + RM_TILE - 1) / RM_TILE;
            while (b < ntile && tmax[b] <= e) ++b;
            if (b < ntile) {
                const uint64_t hi = min(n, (b + 1) * RM_TILE);
                for (j = b * RM_TILE; j < hi; ++j)
                    if (T[j] > e) {
                        res = j;
                        break;
                    }
            }
        }
        nxt[i] = (uint32_t)res;
    }
}

// One doubling round: marked nodes mark their 2^k-th successor, and the
// jump table advances to 2^(k+1).
__global__ void k_rm_mark(const uint32_t* __restrict__ J, uint64_t n2, uint8_t* mark) {
    GRID_STRIDE(i, n2) if (mark[i]) mark[J[i]] = 1;
}
__global__ void k_rm_jump(const uint32_t* __restrict__ J, uint64_t n2, uint32_t* __restrict__ J2) {
    GRID_STRIDE(i, n2) J2[i] = J[J[i]];
}

// Accepted matches: trimmed start max(start, previous accepted end), the
// MapLen byte count of (length - 50), and the text gap before each match.
__global__ void k_rm_accept(const uint64_t* __restrict__ s, const uint64_t* __restrict__ l,
                            const uint64_t* __restrict__ d, uint64_t n, uint64_t* __restrict__ a_start,
                            uint64_t* __restrict__ a_end, uint32_t* __restrict__ len_bytes,
                            uint64_t* __restrict__ gap) {
    GRID_STRIDE(j, n) {
        const uint64_t e = d[j] + l[j];
        const uint64_t prev_end = j ? d[j - 1] + l[j - 1] : 0;
        const uint64_t st = max(d[j], prev_end);
        a_start[j] = st;
        a_end[j] = e;
        uint64_t v = e - st - RM_TARGET;
        uint32_t c = 1;
        while (v >= 128) {
            v >>= 7;
            ++c;
        }
        len_bytes[j] = c;
        gap[j] = st - prev_end + 1;  // text before the match plus its '%'
    }
}

// MapOff (u32 or u64 source positions) and MapLen (byte-frugal) streams
// (compress.cu:975-983, helper.cpp:179-187).
__global__ void k_rm_maps(const uint64_t* __restrict__ s, const uint64_t* __restrict__ a_start,
                          const uint64_t* __restrict__ a_end, const uint32_t* __restrict__ len_off, uint64_t n,
                          bool off32, uint8_t* __restrict__ map_off, uint8_t* __restrict__ map_len) {
    GRID_STRIDE(j, n) {
        const uint64_t v0 = s[j];
        const uint32_t w = off32 ? 4 : 8;
        for (uint32_t b = 0; b < w; ++b) map_off[j * w + b] = (uint8_t)(v0 >> (8 * b));
        uint64_t v = a_end[j] - a_start[j] - RM_TARGET;
        uint8_t* o = map_len + 1 + len_off[j];
        while (v >= 128) {
            *o++ = (uint8_t)(128 + (v & 127));
            v >>= 7;
        }
        *o = (uint8_t)v;
    }
}

// Rewritten text: gap j (text between accepted matches j-1 and j) starts at
// output offset goff[j]; a '%' follows every gap but the last.  Each thread
// writes 64 output bytes.  Characters come from the 2-bit text (binary) or
// pass through the complement table twice (character path, compress.cu:873,
// 953-954).
__global__ void k_rm_text(const uint64_t* __restrict__ dest2, const uint8_t* __restrict__ desta,
                          const uint64_t* __restrict__ a_start, const uint64_t* __restrict__ a_end,
                          const uint64_t* __restrict__ goff, uint64_t n_acc, uint64_t dest_len, uint64_t out_len,
                          uint8_t* __restrict__ out) {
    GRID_STRIDE(c, (out_len + 63) / 64) {
        const uint64_t q0 = c * 64;
        // last gap j with goff[j] <= q0 (goff has n_acc + 1 entries)
        uint64_t lo = 0, hi = n_acc;
        while (lo < hi) {
            const uint64_t mid = (lo + hi + 1) / 2;
            if (goff[mid] <= q0) lo = mid;
            else hi = mid - 1;
        }
        uint64_t j = lo;
        const uint64_t qe = min(out_len, q0 + 64);
        for (uint64_t q = q0; q < qe; ++q) {
            while (j < n_acc && q >= goff[j + 1]) ++j;
            const uint64_t gstart = j ? a_end[j - 1] : 0;
            const uint64_t gend = j < n_acc ? a_start[j] : dest_len;
            const uint64_t k = q - goff[j];
            uint8_t ch;
            if (gstart + k < gend) {
                const uint64_t t = gstart + k;
                ch = dest2 ? (uint8_t)rm_char(base2(dest2, t)) : rm_comp(rm_comp(desta[t]));
            } else {
                ch = '%';
            }
            out[q] = ch;
        }
    }
}

}  // namespace curc
