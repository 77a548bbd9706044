// Device-side 2-bit read helpers for the CURC pseudogenome path on gfx950.
//
// Packed reads use the reference layout (compress.cu:2159-2174): base j of a
// read sits in word j/32, bits [63-2(j%32), 62-2(j%32)], A=0 C=1 G=2 T=3,
// anything else (N included) packs as 0; unused tail bits are zero.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace curc {

constexpr uint32_t INV = 0xFFFFFFFFu;   // constant.hpp:48
constexpr uint16_t INVOFF = 0xFFFFu;    // constant.hpp:49
constexpr uint32_t WAVE = 64;

// MurmurHash3 fmix64 (util.cuh:82-89).
__device__ __forceinline__ uint64_t murmur64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

// One 8-byte load that the compiler keeps apart from loads of the
// neighbouring word.  Two adjacent 8-byte loads are otherwise merged into a
// single 16-byte load at an address that is only 8-byte aligned, and such
// loads have returned wrong data at random on gfx950 (a kernel reading only
// constant arrays gave different results on identical launches); a relaxed
// wavefront-scope atomic load is a plain global_load_dwordx2 that is never
// merged.
__device__ __forceinline__ uint64_t ld64(const uint64_t* p) {
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// The same for 32 bits (u32 arrays read at arbitrary indexes).
__device__ __forceinline__ uint32_t ld32(const uint32_t* p) {
    return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// The same for one byte (byte arrays read at arbitrary offsets).
__device__ __forceinline__ uint8_t ld8(const uint8_t* p) {
    return __hip_atomic_load(const_cast<uint8_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// n (1..32) bases starting at base `off` of the read whose first word is
// `rd`, right-aligned (util.cuh:180-190).  Reads word off/32+1 only when the
// window straddles a word boundary.
__device__ __forceinline__ uint64_t xw(const uint64_t* __restrict__ rd, uint32_t off, uint32_t n) {
    uint32_t o1 = off >> 5;
    uint32_t o2 = (off & 31u) << 1;
    uint64_t left = ld64(rd + o1) << o2;
    if ((64u - o2) >= (n << 1)) return left >> ((32u - n) << 1);
    uint64_t right = (ld64(rd + o1 + 1) >> (62u - o2)) >> 2;
    return (left | right) >> ((32u - n) << 1);
}

// Same window from an array held in registers/LDS (words w[0..]).
template <int MAXW>
__device__ __forceinline__ uint64_t xw_regs(const uint64_t (&w)[MAXW], uint32_t off, uint32_t n) {
    uint32_t o1 = off >> 5;
    uint32_t o2 = (off & 31u) << 1;
    uint64_t left = w[o1] << o2;
    if ((64u - o2) >= (n << 1)) return left >> ((32u - n) << 1);
    uint64_t right = (w[o1 + 1] >> (62u - o2)) >> 2;
    return (left | right) >> ((32u - n) << 1);
}

// Unaligned window over a long 2-bit sequence (util.cuh:191-195); the array
// carries one padding word so the following word is always readable.
__device__ __forceinline__ uint64_t xwu(const uint64_t* __restrict__ bits, uint64_t off, uint32_t n) {
    uint64_t o1 = off >> 5;
    uint32_t o2 = (uint32_t)(off & 31u) << 1;
    uint64_t left = ld64(bits + o1) << o2;
    uint64_t right = (ld64(bits + o1 + 1) >> (62u - o2)) >> 2;
    return (left | right) >> ((32u - n) << 1);
}

__device__ __forceinline__ uint32_t base2(const uint64_t* __restrict__ bits, uint64_t off) {
    return (uint32_t)((ld64(bits + (off >> 5)) >> (((~off) & 31u) << 1)) & 3u);
}

// Reverse complement of a right-aligned k-mer (util.cuh:220-226).
__device__ __forceinline__ uint64_t revcomp(uint64_t km, uint32_t k) {
    km = ~km;
    km = ((km & 0x3333333333333333ULL) << 2) | ((km & 0xCCCCCCCCCCCCCCCCULL) >> 2);
    km = ((km & 0x0F0F0F0F0F0F0F0FULL) << 4) | ((km & 0xF0F0F0F0F0F0F0F0ULL) >> 4);
    km = __builtin_bswap64(km);
    return km >> ((32u - k) << 1);
}

// Number of differing bases between two 2-bit words (util.cuh:228-232).
__device__ __forceinline__ uint32_t mm_words(uint64_t a, uint64_t b) {
    uint64_t x = a ^ b;
    return (uint32_t)__popcll(((x & 0xAAAAAAAAAAAAAAAAULL) >> 1) | (x & 0x5555555555555555ULL));
}

// ASCII -> 2-bit code; non-ACGT -> 4 (util.cuh:148-165).
__device__ __forceinline__ uint32_t enc_base(uint8_t c) {
    // A=65 C=67 G=71 T=84, lowercase +32
    uint32_t u = c & 0xDFu;
    return u == 'A' ? 0u : u == 'C' ? 1u : u == 'G' ? 2u : u == 'T' ? 3u : 4u;
}

__device__ __forceinline__ uint8_t rev_base_ascii(uint8_t b) {
    switch (b) {
        case 'a': return 't'; case 'A': return 'T';
        case 'c': return 'g'; case 'C': return 'G';
        case 'g': return 'c'; case 'G': return 'C';
        case 't': return 'a'; case 'T': return 'A';
        default: return 'N';
    }
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

}  // namespace curc
