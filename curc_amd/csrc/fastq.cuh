// FASTQ records on the device: the record loop of preprocess.cpp:147-199
// (header line, sequence line, '+' line, quality line; only the sequence is
// kept, and its length must be the input's read length, :152-154).
//
// The block's FASTQ bytes are parsed where they land in HBM.  A record is
// four lines, so the sequence of record k starts right after newline 4k
// (numbering newlines from 0).  Pass 1 counts the newlines of each 16 KB
// tile, one workgroup turns the counts into tile offsets, and pass 2 numbers
// every newline and writes the start of each record's sequence line.  The
// pack kernel then reads each read's bytes from there and checks that the
// line holds exactly L bases before its newline.
#pragma once
#include "kernels.cuh"

namespace curc {

constexpr uint32_t FQ_THREAD_BYTES = 64;
constexpr uint32_t FQ_TILE = 256 * FQ_THREAD_BYTES;

// Bit i set for each byte i of d equal to '\n' (exact zero-byte test).
__device__ __forceinline__ uint32_t fq_nl4(uint32_t d) {
    const uint32_t x = d ^ 0x0A0A0A0Au;
    const uint32_t z = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
    return ((z >> 7) & 1u) | ((z >> 14) & 2u) | ((z >> 21) & 4u) | ((z >> 28) & 8u);
}

// Newline mask of the 64 bytes at p (16-byte aligned); bytes at or past
// `end` never count.
__device__ __forceinline__ uint64_t fq_newlines(const uint8_t* p, const uint8_t* end) {
    uint64_t m = 0;
    if (p + FQ_THREAD_BYTES <= end) {
        const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            const uint4 v = q[k];
            const uint64_t b = (uint64_t)(fq_nl4(v.x) | (fq_nl4(v.y) << 4) | (fq_nl4(v.z) << 8) | (fq_nl4(v.w) << 12));
            m |= b << (16 * k);
        }
    } else {
        for (uint32_t j = 0; j < FQ_THREAD_BYTES; ++j)
            if (p + j < end && p[j] == '\n') m |= 1ull << j;
    }
    return m;
}

// Pass 1: newlines per tile of the nb bytes at B (B 16-byte aligned).
__global__ void __launch_bounds__(256) k_fq_count(const uint8_t* __restrict__ B, uint64_t nb,
                                                  uint64_t* __restrict__ tile_cnt) {
    __shared__ uint32_t lds[4];
    const uint64_t off = (uint64_t)blockIdx.x * FQ_TILE + (uint64_t)threadIdx.x * FQ_THREAD_BYTES;
    const uint32_t c = off < nb ? (uint32_t)__popcll(fq_newlines(B + off, B + nb)) : 0u;
    uint32_t tot;
    block_excl_scan256(c, lds, &tot);
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = tot;
}

// Exclusive scan of the n tile counts in place, by one workgroup of 1024
// threads; the total (the block's newline count) goes to *total.
__global__ void __launch_bounds__(1024) k_fq_scan(uint64_t* __restrict__ t, uint64_t n, uint64_t* __restrict__ total) {
    __shared__ uint64_t lds[16];
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint64_t run = 0;
    for (uint64_t base = 0; base < n; base += 1024ull * 4) {
        const uint64_t i0 = base + (uint64_t)threadIdx.x * 4;
        uint64_t v[4];
        uint64_t s = 0;
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) {
            v[r] = (i0 + r < n) ? t[i0 + r] : 0ull;
            s += v[r];
        }
        uint64_t inc = s;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t u = __shfl_up(inc, d, 64);
            if (lane >= (uint32_t)d) inc += u;
        }
        if (lane == 63) lds[w] = inc;
        __syncthreads();
        uint64_t wo = 0, tot = 0;
#pragma unroll
        for (uint32_t i = 0; i < 16; ++i) {
            const uint64_t x = lds[i];
            if (i < w) wo += x;
            tot += x;
        }
        __syncthreads();
        uint64_t ex = run + wo + inc - s;
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) {
            if (i0 + r < n) t[i0 + r] = ex;
            ex += v[r];
        }
        run += tot;
    }
    if (threadIdx.x == 0) *total = run;
}

// Pass 2: numbers the newlines of each tile from its offset and writes, for
// newline 4k, the start of record k's sequence line (the next byte) to
// seq[k * stride + phase] (stride 2 interleaves the two files of a paired
// block: ids 2k and 2k+1, preprocess.cpp:156-192).
__global__ void __launch_bounds__(256) k_fq_lines(const uint8_t* __restrict__ B, uint64_t nb,
                                                  const uint64_t* __restrict__ tile_off, uint64_t* __restrict__ seq,
                                                  uint32_t stride, uint32_t phase) {
    __shared__ uint32_t lds[4];
    const uint64_t off = (uint64_t)blockIdx.x * FQ_TILE + (uint64_t)threadIdx.x * FQ_THREAD_BYTES;
    uint64_t m = off < nb ? fq_newlines(B + off, B + nb) : 0ull;
    uint32_t tot;
    const uint32_t ex = block_excl_scan256((uint32_t)__popcll(m), lds, &tot);
    uint64_t g = tile_off[blockIdx.x] + ex;
    while (m) {
        const uint32_t j = (uint32_t)__builtin_ctzll(m);
        m &= m - 1;
        if ((g & 3u) == 0) seq[(g >> 2) * stride + phase] = off + j + 1;
        ++g;
    }
}

// Format checks of one file's bytes: bit 0 of *flags if it does not start
// with '@', bit 1 if it does not end with a newline.
__global__ void k_fq_ends(const uint8_t* __restrict__ B, uint64_t nb, uint32_t* __restrict__ flags) {
    uint32_t f = 0;
    if (nb == 0 || B[0] != '@') f |= 1u;
    if (nb == 0 || B[nb - 1] != '\n') f |= 2u;
    if (f) atomicOr(flags, f);
}

// Reads of a parsed block: read r is in file (r & paired), its sequence at
// byte start[r] of that file.
struct FqReads {
    const uint8_t* b0;
    const uint8_t* b1;
    uint64_t n0, n1;
    const uint64_t* start;
    uint32_t paired;
    __device__ __forceinline__ uint32_t file(uint64_t r) const { return (uint32_t)(r & paired); }
};

// As k_pack_w, reading read r from its FASTQ sequence line.  Any newline
// among the L bytes, or a byte other than a newline right after them, means
// the line does not hold L bases ("read length isn't fix",
// preprocess.cpp:154): *err is set and the block fails.
__global__ void __launch_bounds__(256) k_pack_fq(FqReads in, uint64_t n, uint32_t L, uint32_t W,
                                                 uint64_t* __restrict__ packed, uint8_t* __restrict__ isN,
                                                 uint32_t* __restrict__ err) {
    GRID_STRIDE(t, n * W) {
        const uint64_t r = t / W;
        const uint32_t w = (uint32_t)(t - r * W);
        const uint32_t f = in.file(r);
        const uint8_t* fb = f ? in.b1 : in.b0;
        const uint64_t fn = f ? in.n1 : in.n0;
        const uint64_t s = in.start[r];
        const uint32_t b0 = 32u * w;
        const uint32_t nb = min(32u, L - b0);
        uint32_t a[8];
        load_bytes32(fb + s + b0, nb, fb, fb + fn, a);
        uint64_t v = 0;
        uint32_t hasN = 0, hasNL = 0;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) {
            const uint32_t valid = 4 * k < nb ? min(4u, nb - 4 * k) : 0u;
            const uint32_t keep = valid >= 4 ? 0xFFFFFFFFu : ((1u << (8 * valid)) - 1u);
            const uint32_t x = (a[k] ^ 0x4E4E4E4Eu) | ~keep;
            hasN |= (x - 0x01010101u) & ~x & 0x80808080u;
            hasNL |= fq_nl4(a[k]) & (valid >= 4 ? 0xFu : ((1u << valid) - 1u));
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
        bool bad = hasNL != 0 || s + b0 + nb > fn;
        if (w == W - 1 && !bad) bad = s + L >= fn || fb[s + L] != '\n';
        if (bad) atomicOr(err, 1u);
        packed[t] = v;
        if (isN && hasN) {
            const uintptr_t fp = (uintptr_t)(isN + r);
            atomicOr(reinterpret_cast<uint32_t*>(fp & ~(uintptr_t)3), 1u << (8 * (fp & 3)));
        }
    }
}

// The N-reads' sequences (ids in N-index order) from the FASTQ lines.
__global__ void k_gather_ascii_fq(FqReads in, const uint32_t* __restrict__ ids, uint64_t n, uint32_t L,
                                  uint8_t* __restrict__ out) {
    GRID_STRIDE(b, n * L) {
        const uint64_t r = b / L, j = b - r * L;
        const uint32_t id = ids[r];
        const uint8_t* fb = in.file(id) ? in.b1 : in.b0;
        out[b] = fb[in.start[id] + j];
    }
}

}  // namespace curc
