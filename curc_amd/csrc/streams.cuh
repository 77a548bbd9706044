// Record sort and stream builders of one block (compress.cu:1730-2044),
// computed where the records already are.  The output is CURC's pre-entropy
// parity surface: read order, positions, strands and mismatch lists.
//
// Records are HQ records followed by map records (compress.cu:1737-1742),
// stably sorted by position, or by global read id when order is preserved.
// Per sorted record r the builders write its strand character, mismatch
// count and position delta; a record with m mismatches writes m symbols to
// the base stream and width(m) values to the offset stream for m:
// width = 2 (m == 1), 3 (m == 2), m + 2 (m > 2).
#pragma once
#include "kernels.cuh"

namespace curc {

__host__ __device__ __forceinline__ uint32_t mis_width(uint32_t m) { return m == 1 ? 2u : m == 2 ? 3u : m + 2u; }

struct RecSrc {
    const uint32_t* hq_id;    // global ids (HQ records: strand 0, not N)
    const uint64_t* hq_pos;
    uint32_t n_hq;
    const uint32_t* map_id;   // global id, or N index when is_contain_N
    const uint64_t* map_word; // is_contain_N | strand << 1 | pos << 2
    const uint8_t* map_mis;
    const uint32_t* Nids;     // N index -> global id
    __device__ void get(uint32_t r, uint32_t& id, uint64_t& word, uint32_t& m) const {
        if (r < n_hq) {
            id = hq_id[r];
            word = hq_pos[r] << 2;
            m = 0;
        } else {
            id = map_id[r - n_hq];
            word = map_word[r - n_hq];
            m = map_mis[r - n_hq];
        }
    }
    __device__ uint32_t gid(uint32_t id, uint64_t word) const { return (word & 1) ? Nids[id] : id; }
};

// Sort keys: position (order not preserved) or global id (preserved).
__global__ void k_rec_keys(RecSrc src, uint32_t n, bool preserve, uint64_t* __restrict__ keys,
                           uint32_t* __restrict__ vals) {
    GRID_STRIDE(r, n) {
        uint32_t id, m;
        uint64_t w;
        src.get((uint32_t)r, id, w, m);
        keys[r] = preserve ? (uint64_t)src.gid(id, w) : (w >> 2);
        vals[r] = (uint32_t)r;
    }
}

struct StreamOut {
    uint32_t* rid;        // sorted records
    uint64_t* rword;
    uint8_t* strand;      // '0' / '1'
    uint8_t* mis_count;
    uint16_t* read_off;   // order not preserved
    uint32_t* pe_order;   // paired, order not preserved
    uint64_t* id_pos;     // order preserved
    uint32_t* mkey;       // mismatch count per sorted record (for grouping by m)
    uint32_t* midx;       // sorted record index
};

__global__ void k_rec_streams(RecSrc src, const uint32_t* __restrict__ order, uint32_t n, bool preserve,
                              bool paired, StreamOut o) {
    GRID_STRIDE(i, n) {
        uint32_t id, m;
        uint64_t w;
        src.get(order[i], id, w, m);
        o.rid[i] = id;
        o.rword[i] = w;
        o.strand[i] = (uint8_t)('0' + ((w >> 1) & 1));
        o.mis_count[i] = (uint8_t)m;
        o.mkey[i] = m;
        o.midx[i] = (uint32_t)i;
        const uint32_t g = src.gid(id, w);
        if (preserve) {
            o.id_pos[g] = w >> 2;
        } else {
            if (paired) o.pe_order[i] = g;
            uint64_t prev_pos = 0;
            if (i > 0) {
                uint32_t pid, pm;
                uint64_t pw;
                src.get(order[i - 1], pid, pw, pm);
                prev_pos = pw >> 2;
            }
            o.read_off[i] = (uint16_t)(i == 0 ? 0 : ((w >> 2) - prev_pos));
        }
    }
}

// Read base p of the sequence as it lies on the pseudogenome (strand 1 reads
// are reverse complemented), as a character (compress.cu:1820-1845).
__device__ __forceinline__ uint8_t placed_base(const uint64_t* R, uint32_t W, const uint8_t* nascii, uint32_t L,
                                               uint32_t id, uint64_t word, uint32_t p) {
    const bool isN = word & 1, rc = (word >> 1) & 1;
    const uint32_t bo = rc ? L - 1 - p : p;
    uint8_t b = isN ? nascii[(uint64_t)id * L + bo] : base_char(base2(R + (uint64_t)id * W, bo));
    return rc ? rev_base_ascii(b) : b;
}

__device__ __forceinline__ uint32_t ctx_index(uint8_t c) {  // contexts A, T, C, G (compress.cu:1914-1917)
    return c == 'A' ? 0u : c == 'T' ? 1u : c == 'C' ? 2u : 3u;
}
__device__ __forceinline__ uint32_t sym_index(uint8_t c) {  // read bases A, C, G, N, T
    return c == 'A' ? 0u : c == 'C' ? 1u : c == 'G' ? 2u : c == 'T' ? 4u : 3u;
}

struct MisArgs {
    const uint64_t* R;
    uint32_t W, L;
    const uint8_t* nascii;
    const uint64_t* pg;       // 2-bit HQ pseudogenome
    const uint32_t* rid;      // sorted records
    const uint64_t* rword;
    const uint8_t* mis_count;
    const uint32_t* base_idx; // exclusive scan of mismatch counts
    const uint32_t* off_idx;  // start of the record's values in the offset stream
    uint8_t* pair;            // per mismatch: ctx_index * 5 + sym_index
    uint16_t* off_stream;     // all offset streams, concatenated by m
    uint32_t* hist;           // [4][5] context/symbol counts
};

// Recomputes each mapped record's mismatches against the pseudogenome and
// writes its symbols and its offset-stream values (compress.cu:1815-1886).
__global__ void __launch_bounds__(256) k_rec_mismatch(MisArgs a, uint32_t n) {
    __shared__ uint32_t h[20];
    if (threadIdx.x < 20) h[threadIdx.x] = 0;
    __syncthreads();
    GRID_STRIDE(i, n) {
        const uint32_t m = a.mis_count[i];
        if (m == 0) continue;
        const uint32_t id = a.rid[i];
        const uint64_t w = a.rword[i];
        const uint64_t pos = w >> 2;
        uint32_t k = a.base_idx[i];
        uint32_t first = 0, last = 0, cnt = 0, min_gap = 0xFFFFFFFFu, prevp = 0;
        for (uint32_t p = 0; p < a.L; ++p) {
            uint8_t rb = placed_base(a.R, a.W, a.nascii, a.L, id, w, p);
            uint8_t gb = base_char(base2(a.pg, pos + p));
            if (rb == gb) continue;
            uint32_t code = ctx_index(gb) * 5 + sym_index(rb);
            a.pair[k++] = (uint8_t)code;
            atomicAdd(&h[code], 1u);
            if (cnt == 0) first = p;
            else min_gap = min(min_gap, p - prevp);
            prevp = p;
            last = p;
            ++cnt;
        }
        uint16_t* S = a.off_stream + a.off_idx[i];
        uint32_t q = 0;
        const bool l2r = first < (a.L - 1 - last);
        S[q++] = l2r ? 0 : 1;
        S[q++] = (uint16_t)(l2r ? first : (a.L - 1 - last));
        if (m == 2) {
            S[q++] = (uint16_t)(last - first - 1);
        } else if (m > 2) {
            S[q++] = (uint16_t)(min_gap - 1);
            // gaps in emission order: ascending for l2r, descending otherwise
            uint32_t pp = l2r ? 0xFFFFFFFFu : a.L;
            for (uint32_t t = 0; t < a.L; ++t) {
                uint32_t p = l2r ? t : a.L - 1 - t;
                uint8_t rb = placed_base(a.R, a.W, a.nascii, a.L, id, w, p);
                uint8_t gb = base_char(base2(a.pg, pos + p));
                if (rb == gb) continue;
                if (pp != 0xFFFFFFFFu && pp != a.L) S[q++] = (uint16_t)((l2r ? p - pp : pp - p) - min_gap);
                pp = p;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 20 && h[threadIdx.x]) atomicAdd(&a.hist[threadIdx.x], h[threadIdx.x]);
}

// Offset-stream start per record: base of its m-stream plus width(m) times
// its rank among records with m mismatches (records grouped by a stable sort
// on m, so rank = position in the group).
__global__ void k_rec_offidx(const uint32_t* __restrict__ mk_sorted, const uint32_t* __restrict__ idx_sorted,
                             uint32_t n, const uint32_t* __restrict__ group_start, const uint64_t* __restrict__ mbase,
                             uint32_t* __restrict__ off_idx) {
    GRID_STRIDE(q, n) {
        uint32_t m = mk_sorted[q];
        if (m == 0) continue;
        uint32_t rank = (uint32_t)q - group_start[m];
        off_idx[idx_sorted[q]] = (uint32_t)(mbase[m] + (uint64_t)mis_width(m) * rank);
    }
}

__global__ void k_rec_recode(const uint8_t* __restrict__ pair, uint64_t n, const uint8_t* __restrict__ rank,
                             uint8_t* __restrict__ sym) {
    GRID_STRIDE(i, n) sym[i] = rank[pair[i]];
}

// LQ / N-LQ records: offsets L - (pos - previous pos), pair order, id ->
// joined-pseudogenome position (compress.cu:1968-2001).
__global__ void k_lq_streams(const uint32_t* __restrict__ rid, const uint64_t* __restrict__ pos, uint32_t n,
                             const uint32_t* __restrict__ idmap, uint32_t L, bool preserve, bool paired,
                             uint64_t pg_base, uint16_t* __restrict__ off, uint32_t* __restrict__ pe_order,
                             uint64_t* __restrict__ id_pos) {
    GRID_STRIDE(i, n) {
        uint32_t g = idmap ? idmap[rid[i]] : rid[i];
        if (preserve) {
            id_pos[g] = pg_base + pos[i];
        } else {
            uint64_t pp = i ? pos[i - 1] : 0;
            off[i] = (uint16_t)(L - (pos[i] - pp));
            if (paired) pe_order[i] = g;
        }
    }
}

}  // namespace curc
