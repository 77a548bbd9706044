// Pseudogenome assembly on the GPU (compress.cu:107-165, RefRecord::assembly).
//
// The reference walks the link graph on the host: for each read of the visit
// list not yet placed, it walks back to the chain head (or around a cycle,
// which it cuts after the first largest-offset read met from next[rid]) and
// emits the chain.  Here the same output comes from list ranking:
//  1. nodes are renamed to their rank in the visit list;
//  2. pointer jumping along prev gives, per node, the bases emitted before it
//     in its chain, its index in the chain, its head and the minimum rank in
//     its chain (or, for a cycle, in the whole cycle);
//  3. each cycle is ranked from its member of minimum rank (the read the
//     reference visits first), the reference's cut is found from those
//     distances, and the cycle is ranked again as a path;
//  4. chains are emitted in order of their minimum rank, which is the order in
//     which the reference's visit loop first touches them.
#pragma once
#include "kernels.cuh"

namespace curc {

// 32-byte aligned so that the compiler's 16-byte loads of it are aligned.
struct alignas(32) JumpState {
    uint32_t p;  // jump pointer (rank) or INV
    uint32_t c;  // nodes in the window
    uint32_t h;  // last node of the window (chain head once p == INV)
    uint32_t m;  // minimum rank in the window
    uint64_t s;  // bases emitted by the window's nodes
};

__global__ void k_asm_rank(const uint32_t* __restrict__ V, uint64_t n, uint32_t* __restrict__ rank) {
    GRID_STRIDE(j, n) rank[V[j]] = (uint32_t)j;
}

__global__ void k_asm_links(const uint32_t* __restrict__ V, uint64_t n, const uint32_t* __restrict__ next,
                            const uint32_t* __restrict__ prev, const uint16_t* __restrict__ offset,
                            const uint32_t* __restrict__ rank, uint32_t* __restrict__ nxtR,
                            uint32_t* __restrict__ prvR, uint16_t* __restrict__ offR) {
    GRID_STRIDE(j, n) {
        uint32_t i = V[j];
        uint32_t nx = next[i], pv = prev[i];
        nxtR[j] = (nx == INV) ? INV : rank[nx];
        prvR[j] = (pv == INV) ? INV : rank[pv];
        offR[j] = offset[i];
    }
}

// A node emits offset[i] bases, or the whole read if it ends its chain.
__device__ __forceinline__ uint32_t emitted(const uint32_t* nxtR, const uint16_t* offR, uint64_t j, uint32_t L) {
    return nxtR[j] == INV ? L : (uint32_t)offR[j];
}

__global__ void k_asm_init(const uint32_t* __restrict__ list, uint64_t n, const uint32_t* __restrict__ nxtR,
                           const uint32_t* __restrict__ prvR, const uint16_t* __restrict__ offR, uint32_t L,
                           JumpState* __restrict__ st) {
    GRID_STRIDE(q, n) {
        uint32_t j = list ? list[q] : (uint32_t)q;
        JumpState a;
        a.p = prvR[j];
        a.c = 1;
        a.h = j;
        a.m = j;
        a.s = emitted(nxtR, offR, j, L);
        st[j] = a;
    }
}

__global__ void k_asm_round(const uint32_t* __restrict__ list, uint64_t n, const JumpState* __restrict__ in,
                            JumpState* __restrict__ out) {
    GRID_STRIDE(q, n) {
        uint32_t j = list ? list[q] : (uint32_t)q;
        JumpState a = in[j];
        if (a.p != INV) {
            JumpState b = in[a.p];
            a.s += b.s;
            a.c += b.c;
            a.h = b.h;
            a.m = min(a.m, b.m);
            a.p = b.p;
        }
        out[j] = a;
    }
}

// List ranking by a ruling set.  The jump rounds over all n nodes cost
// log2(n) passes over the states; instead the chain heads and about one
// node in 8 (a multiplicative hash of the rank: fixed, not random) are
// rulers, each ruler walks its sublist along next up to the next ruler, and
// only the rulers are ranked by jump rounds, on the sums of the sublists
// before them.  A node's final state (the sums from its chain head to
// itself, the head, the minimum rank) is its ruler's exclusive prefix plus
// its own local sums, the same values the rounds over all nodes give.  A
// cycle holding a ruler is ranked around by its rulers (its rulers' pointers
// stay live and their windows wrap the cycle); a cycle without one is
// ranked by the rounds over its members alone.
constexpr uint32_t ASM_RULER_SHIFT = 29;  // rulers: top 3 bits of the rank's hash clear
__device__ __forceinline__ bool asm_ruler_hash(uint32_t j) { return ((j * 0x9E3779B1u) >> ASM_RULER_SHIFT) == 0; }

// The rulers in rank order (a stable compaction): rlist, their index ridx,
// and the state of a chain head (nothing before it; a later ruler's state is
// written by the walk of the ruler before it).
struct PredRuler {
    const uint32_t* prvR;
    __device__ bool operator()(uint32_t j) const { return prvR[j] == INV || asm_ruler_hash(j); }
};
struct EmitRuler {
    const uint32_t* prvR;
    uint32_t* rlist;
    uint32_t* ridx;
    JumpState* rs;
    __device__ void operator()(uint32_t j, uint32_t q) const {
        rlist[q] = j;
        ridx[j] = q;
        if (prvR[j] == INV) {
            JumpState a;
            a.p = INV;
            a.c = 0;
            a.h = j;
            a.m = INV;
            a.s = 0;
            rs[q] = a;
        }
    }
};

// Ruler q walks its sublist: each node gets its ruler and the sums from the
// ruler to itself (loc); the next ruler gets the sums of the whole sublist
// and q as its predecessor.
__global__ void k_asm_walk(uint64_t nr, const uint32_t* __restrict__ rlist, const uint32_t* __restrict__ nxtR,
                           const uint16_t* __restrict__ offR, uint32_t L, const uint32_t* __restrict__ ridx,
                           uint32_t* __restrict__ rul, JumpState* __restrict__ loc, JumpState* __restrict__ rs) {
    GRID_STRIDE(q, nr) {
        uint32_t x = rlist[q];
        JumpState a;
        a.p = (uint32_t)q;
        a.c = 0;
        a.h = 0;
        a.m = INV;
        a.s = 0;
        for (;;) {
            const uint32_t y = nxtR[x];
            a.c += 1;
            a.s += y == INV ? L : (uint32_t)offR[x];
            a.m = min(a.m, x);
            rul[x] = (uint32_t)q;
            loc[x] = a;
            if (y == INV) break;
            // y has a predecessor (x), so it is a ruler by its hash alone
            if (asm_ruler_hash(y)) {
                rs[ridx[y]] = a;
                break;
            }
            x = y;
        }
    }
}

// Final states from the rulers' ranks: nodes of a cycle without a ruler are
// listed for the rounds over their members.
__global__ void k_asm_final(uint64_t n, const uint32_t* __restrict__ rul, const JumpState* __restrict__ loc,
                            const JumpState* __restrict__ rs, JumpState* __restrict__ st,
                            uint32_t* __restrict__ rest, uint32_t* __restrict__ n_rest) {
    GRID_STRIDE(j, n) {
        const uint32_t q = rul[j];
        const uint32_t slot = wave_append(n_rest, q == INV);
        if (q == INV) {
            rest[slot] = (uint32_t)j;
            continue;
        }
        const JumpState r = rs[q], l = loc[j];
        JumpState a;
        a.p = r.p;
        a.c = r.c + l.c;
        a.h = r.h;
        a.m = min(r.m, l.m);
        a.s = r.s + l.s;
        st[j] = a;
    }
}

// Cycle nodes keep a live jump pointer after enough rounds; their window
// holds the whole cycle, so m is the cycle's minimum rank j0, the member the
// reference's visit loop meets first.  The reference cuts the cycle after
// the first read of largest offset met walking next from j0 (compress.cu:
// 125-145): the walk visits next[j0] (distance 1), ..., j0 last (distance
// len).  Here that cut is found in parallel:
//  1. k_asm_cyc_list lists the cycle members and records their cycle (cid);
//  2. k_asm_cyc_open opens each cycle at j0 (no predecessor) and the jump
//     rounds rank the members as a path: c = 1 + distance from j0;
//  3. k_asm_cyc_best takes, per cycle, the largest (offset, -distance) with
//     j0 at the largest distance, which is that first read of largest offset;
//  4. k_asm_cyc_cut cuts after it; the cycles are then ranked again as paths.
__global__ void k_asm_cyc_list(uint64_t n, const JumpState* __restrict__ st, uint32_t* __restrict__ members,
                               uint32_t* __restrict__ n_members, uint32_t* __restrict__ cid,
                               unsigned long long* __restrict__ best) {
    GRID_STRIDE(j, n) {
        const JumpState a = st[j];
        const uint32_t slot = wave_append(n_members, a.p != INV);
        if (a.p == INV) continue;
        members[slot] = (uint32_t)j;
        cid[j] = a.m;
        if (a.m == (uint32_t)j) best[j] = 0ull;
    }
}

__global__ void k_asm_cyc_open(const uint32_t* __restrict__ members, uint64_t nm, const uint32_t* __restrict__ cid,
                               const uint32_t* __restrict__ prvR, JumpState* __restrict__ st) {
    GRID_STRIDE(q, nm) {
        const uint32_t j = members[q];
        JumpState a;
        a.p = cid[j] == j ? INV : prvR[j];
        a.c = 1;
        a.h = j;
        a.m = j;
        a.s = 0;
        st[j] = a;
    }
}

// Order key of a member in the walk from its cycle's j0: larger offset
// first, then smaller distance (never 0: the key of every member is > 0).
__device__ __forceinline__ unsigned long long cyc_key(uint32_t j, const uint32_t* cid, const JumpState* st,
                                                      const uint16_t* offR) {
    const uint32_t d = cid[j] == j ? 0xFFFFFFFEu : st[j].c - 1;
    return ((unsigned long long)offR[j] << 32) | (0xFFFFFFFFu - d);
}

__global__ void k_asm_cyc_best(const uint32_t* __restrict__ members, uint64_t nm, const uint32_t* __restrict__ cid,
                               const JumpState* __restrict__ st, const uint16_t* __restrict__ offR,
                               unsigned long long* __restrict__ best) {
    GRID_STRIDE(q, nm) {
        const uint32_t j = members[q];
        atomicMax(&best[cid[j]], cyc_key(j, cid, st, offR));
    }
}

__global__ void k_asm_cyc_cut(const uint32_t* __restrict__ members, uint64_t nm, const uint32_t* __restrict__ cid,
                              const JumpState* __restrict__ st, const uint16_t* __restrict__ offR,
                              const unsigned long long* __restrict__ best, uint32_t* __restrict__ nxtR,
                              uint32_t* __restrict__ prvR) {
    GRID_STRIDE(q, nm) {
        const uint32_t j = members[q];
        if (cyc_key(j, cid, st, offR) != best[cid[j]]) continue;
        const uint32_t hd = nxtR[j];
        nxtR[j] = INV;
        prvR[hd] = INV;
    }
}

// Chain tails publish their chain's key (minimum rank), length and size at
// the chain head.
__global__ void k_asm_tails(uint64_t n, const JumpState* __restrict__ st, const uint32_t* __restrict__ nxtR,
                            uint32_t* __restrict__ ckey, uint64_t* __restrict__ clen, uint32_t* __restrict__ ccnt) {
    GRID_STRIDE(j, n) {
        if (nxtR[j] != INV) continue;
        JumpState a = st[j];
        ckey[a.h] = a.m;
        clen[a.h] = a.s;
        ccnt[a.h] = a.c;
    }
}

struct PredHead {
    const uint32_t* prvR;
    __device__ bool operator()(uint32_t j) const { return prvR[j] == INV; }
};
struct EmitHead {
    const uint32_t* ckey;
    uint32_t* keys;
    uint32_t* heads;
    __device__ void operator()(uint32_t j, uint32_t q) const {
        keys[q] = ckey[j];
        heads[q] = j;
    }
};

__global__ void k_asm_gather_sizes(const uint32_t* __restrict__ heads, uint64_t nh, const uint64_t* __restrict__ clen,
                                   const uint32_t* __restrict__ ccnt, uint64_t* __restrict__ olen,
                                   uint32_t* __restrict__ ocnt) {
    GRID_STRIDE(q, nh) {
        uint32_t h = heads[q];
        olen[q] = clen[h];
        ocnt[q] = ccnt[h];
    }
}

__global__ void k_asm_scatter_bases(const uint32_t* __restrict__ heads, uint64_t nh, const uint64_t* __restrict__ olen,
                                    const uint32_t* __restrict__ ocnt, uint64_t* __restrict__ cpos,
                                    uint32_t* __restrict__ crec) {
    GRID_STRIDE(q, nh) {
        uint32_t h = heads[q];
        cpos[h] = olen[q];
        crec[h] = ocnt[q];
    }
}

__global__ void k_asm_records(const uint32_t* __restrict__ V, uint64_t n, const JumpState* __restrict__ st,
                              const uint32_t* __restrict__ nxtR, const uint16_t* __restrict__ offR, uint32_t L,
                              const uint64_t* __restrict__ cpos, const uint32_t* __restrict__ crec,
                              uint32_t* __restrict__ rid, uint64_t* __restrict__ rpos, uint32_t* __restrict__ rnb) {
    GRID_STRIDE(j, n) {
        JumpState a = st[j];
        uint32_t e = emitted(nxtR, offR, j, L);
        uint32_t idx = crec[a.h] + a.c - 1;
        rid[idx] = V[j];
        rpos[idx] = cpos[a.h] + a.s - e;
        rnb[idx] = e;
    }
}

}  // namespace curc
