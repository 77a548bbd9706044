// Read reconstruction of curc -d on the device (decompress.cu:909-1005;
// preserve_order_process, :472-577).  Every read is a slice of one of the
// three pseudogenome texts; a read of the HQ text (HQ and mapped records)
// then gets its mismatch symbols written in and, on the reverse strand, is
// reverse complemented.
//
// Inputs are the block's decoded streams: the texts, per-record strand and
// mismatch count, mismatch symbols, the mismatch-offset streams grouped by
// count, and either the read offsets (order not preserved) or each read's
// position in the joined texts (order preserved).  Offsets are rebuilt
// first into one flat array in record order (as the reference's
// mismatch_offset_decompress does, :357-409); then one workgroup rebuilds 64
// output lines in LDS and writes them out contiguously.
#pragma once
#include "kernels.cuh"

namespace curc {

constexpr uint32_t DEC_LINES = 64;  // output lines per workgroup

// Values whose inclusive prefix sums are the read positions of the three
// parts (decompress_read_off_stream, :67-82): HQ read_off as is, LQ and
// N-LQ offsets as L - off (u16 arithmetic, as the reference's).
__global__ void k_dec_pos_values(const uint16_t* __restrict__ read_off, uint64_t nh, const uint16_t* __restrict__ lq_off,
                                 uint64_t nl, const uint16_t* __restrict__ nlq_off, uint64_t nn, uint32_t L,
                                 uint64_t* __restrict__ v) {
    GRID_STRIDE(e, nh + nl + nn) {
        uint16_t x;
        if (e < nh) x = read_off[e];
        else if (e < nh + nl) x = (uint16_t)(L - lq_off[e - nh]);
        else x = (uint16_t)(L - nlq_off[e - nh - nl]);
        v[e] = x;
    }
}

__global__ void k_dec_u8_to_u64(const uint8_t* __restrict__ a, uint64_t n, uint64_t* __restrict__ v) {
    GRID_STRIDE(i, n) v[i] = a[i];
}

// 1 for reads whose joined position lies in the HQ text (order preserved).
__global__ void k_dec_hq_mask(const uint64_t* __restrict__ id_pos, uint64_t n, uint64_t hq_len,
                              uint64_t* __restrict__ v) {
    GRID_STRIDE(i, n) v[i] = id_pos[i] < hq_len ? 1u : 0u;
}

__global__ void k_dec_count_keys(const uint8_t* __restrict__ cnt, uint64_t n, uint32_t* __restrict__ key,
                                 uint32_t* __restrict__ idx) {
    GRID_STRIDE(i, n) {
        key[i] = cnt[i];
        idx[i] = (uint32_t)i;
    }
}

// Start of each count's run in the records sorted (stably) by count.
__global__ void k_dec_group_start(const uint32_t* __restrict__ key, uint64_t n, uint64_t* __restrict__ start) {
    GRID_STRIDE(s, n) {
        if (s == 0 || key[s] != key[s - 1]) start[key[s]] = s;
    }
}

// Offset-stream row width for m mismatches (flag, first offset, then one
// gap for m = 2, or the minimum gap and m - 1 gaps for m >= 3;
// compress.cu:1848-1886).
__device__ __forceinline__ uint32_t dec_row_width(uint32_t m) { return m == 1 ? 2u : (m == 2 ? 3u : m + 2u); }

// Record r (the s-th by count) with m mismatches is the g-th row of the
// stream for m; its offsets go to flat[bidx[r] .. + m) in ascending order
// (mismatch_offset_decompress, decompress.cu:382-406).
__global__ void k_dec_offsets(const uint32_t* __restrict__ key, const uint32_t* __restrict__ idx, uint64_t n,
                              const uint64_t* __restrict__ gstart, const uint16_t* __restrict__ streams,
                              const uint64_t* __restrict__ sstart, const uint64_t* __restrict__ bidx, uint32_t L,
                              uint32_t max_mm, uint16_t* __restrict__ flat, uint32_t* __restrict__ err) {
    GRID_STRIDE(s, n) {
        const uint32_t m = key[s];
        if (m == 0) continue;
        if (m > max_mm) {
            atomicOr(err, 1u);
            continue;
        }
        const uint32_t w = dec_row_width(m);
        const uint64_t g = s - gstart[m];
        const uint64_t a = sstart[m] + g * w;
        if (a + w > sstart[m + 1]) {
            atomicOr(err, 1u);
            continue;
        }
        const uint16_t* row = streams + a;
        uint16_t* out = flat + bidx[idx[s]];
        const uint32_t flag = row[0], first = row[1];
        bool bad = false;
        if (flag == 0) {
            int64_t p = first;
            out[0] = (uint16_t)p;
            bad |= p >= (int64_t)L;
            if (m == 2) {
                p += (int64_t)row[2] + 1;
                out[1] = (uint16_t)p;
                bad |= p >= (int64_t)L;
            } else if (m > 2) {
                const int64_t mn = (int64_t)row[2] + 1;
                for (uint32_t k = 1; k < m; ++k) {
                    p += (int64_t)row[2 + k] + mn;
                    out[k] = (uint16_t)p;
                    bad |= p >= (int64_t)L;
                }
            }
        } else {
            int64_t p = (int64_t)L - 1 - (int64_t)first;
            out[m - 1] = (uint16_t)p;
            bad |= p < 0;
            if (m == 2) {
                p -= (int64_t)row[2] + 1;
                out[0] = (uint16_t)p;
                bad |= p < 0;
            } else if (m > 2) {
                const int64_t mn = (int64_t)row[2] + 1;
                for (uint32_t k = m - 1; k > 0; --k) {
                    p -= (int64_t)row[2 + (m - k)] + mn;
                    out[k - 1] = (uint16_t)p;
                    bad |= p < 0;
                }
            }
        }
        if (bad) atomicOr(err, 1u);
    }
}

// Complement of the reverse strand (reverse_base, decompress.cu:444-456).
__device__ __forceinline__ uint8_t dec_comp(uint8_t b) {
    switch (b) {
        case 'a': return 't'; case 'A': return 'T';
        case 'c': return 'g'; case 'C': return 'G';
        case 'g': return 'c'; case 'G': return 'C';
        case 't': return 'a'; case 'T': return 'A';
        default: return 'N';
    }
}

struct DecArgs {
    const uint8_t* text;       // joined HQ | LQ | N-LQ text
    uint64_t hq_len, lq_len, nlq_len;
    uint64_t n_reads;          // output lines (both files when paired)
    uint32_t L;
    uint32_t paired;
    // order not preserved: emission index e of output slot i (pe_slot, paired only) ...
    const uint32_t* pe_slot;
    uint64_t nh, nl;           // HQ/mapped records, LQ records (N-LQ: the rest)
    const uint64_t* pos_ex;    // ... exclusive prefix of the position values, and the values
    const uint64_t* pos_v;
    // order preserved: joined position of read i, and its HQ record index
    const uint64_t* id_pos;
    const uint64_t* hq_idx;
    // HQ records
    const char* strand;
    const uint8_t* mis_count;
    const uint64_t* bidx;      // first mismatch of each record
    const uint16_t* flat_off;
    const uint8_t* mis_base;
    uint64_t n_mis;
    const uint8_t* table;      // 16 bytes: rows for contexts A, T, C, G
    uint8_t* out1;
    uint8_t* out2;
    uint32_t* err;
};

// One workgroup per DEC_LINES output slots (slot i: line i of out1, or when
// paired line i/2 of out1 (even i) / out2 (odd i)).
template <bool PRESERVE>
__global__ void __launch_bounds__(256) k_dec_reads(DecArgs a) {
    extern __shared__ uint8_t lines[];  // DEC_LINES x (L + 1)
    __shared__ uint64_t src[DEC_LINES];
    __shared__ int64_t rec[DEC_LINES];
    __shared__ uint8_t rc[DEC_LINES];
    const uint32_t L = a.L, LL = L + 1;
    const uint64_t i0 = (uint64_t)blockIdx.x * DEC_LINES;
    const uint32_t nl = (uint32_t)min<uint64_t>(DEC_LINES, a.n_reads - i0);
    if (threadIdx.x < nl) {
        const uint64_t i = i0 + threadIdx.x;
        uint64_t p, lim;  // position in the joined text, end of its part
        int64_t r = -1;
        if (PRESERVE) {
            const uint64_t q = a.id_pos[i];
            if (q < a.hq_len) {
                r = (int64_t)a.hq_idx[i];
                lim = a.hq_len;
            } else if (q < a.hq_len + a.lq_len) {
                lim = a.hq_len + a.lq_len;
            } else {
                lim = a.hq_len + a.lq_len + a.nlq_len;
            }
            p = q;
        } else {
            uint64_t e = a.paired ? a.pe_slot[i] : i;
            if (e >= a.n_reads) {
                atomicOr(a.err, 1u);
                e = 0;
            }
            const uint64_t rel = a.pos_ex[e] + a.pos_v[e];
            if (e < a.nh) {
                r = (int64_t)e;
                p = rel;
                lim = a.hq_len;
            } else if (e < a.nh + a.nl) {
                p = a.hq_len + rel;
                lim = a.hq_len + a.lq_len;
            } else {
                p = a.hq_len + a.lq_len + rel;
                lim = a.hq_len + a.lq_len + a.nlq_len;
            }
        }
        if (p + L > lim) {
            atomicOr(a.err, 1u);
            p = 0;
            r = -1;
        }
        src[threadIdx.x] = p;
        rec[threadIdx.x] = r;
        rc[threadIdx.x] = (r >= 0 && a.strand[r] == '1') ? 1 : 0;
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < nl * L; t += blockDim.x) {
        const uint32_t r = t / L, k = t - r * L;
        lines[r * LL + k] = a.text[src[r] + k];
    }
    __syncthreads();
    if (threadIdx.x < nl && rec[threadIdx.x] >= 0) {
        // mismatch symbols into the forward slice (decompress.cu:955-965)
        const uint64_t r = (uint64_t)rec[threadIdx.x];
        const uint32_t m = a.mis_count[r];
        const uint64_t b = a.bidx[r];
        if (b + m > a.n_mis) {
            atomicOr(a.err, 1u);
        } else {
            uint8_t* ln = lines + threadIdx.x * LL;
            for (uint32_t k = 0; k < m; ++k) {
                const uint32_t o = a.flat_off[b + k];
                const uint32_t sym = a.mis_base[b + k];
                if (o >= L || sym > 3) {
                    atomicOr(a.err, 1u);
                    break;
                }
                const uint8_t c = ln[o];
                const uint32_t row = c == 'T' ? 1u : c == 'C' ? 2u : c == 'G' ? 3u : 0u;
                ln[o] = a.table[row * 4 + sym];
            }
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < nl * LL; t += blockDim.x) {
        const uint32_t r = t / LL, k = t - r * LL;
        uint8_t v;
        if (k == L) v = '\n';
        else if (rc[r]) v = dec_comp(lines[r * LL + (L - 1 - k)]);
        else v = lines[r * LL + k];
        const uint64_t i = i0 + r;
        if (a.paired) {
            uint8_t* o = (i & 1) ? a.out2 : a.out1;
            o[(i >> 1) * LL + k] = v;
        } else {
            a.out1[i * LL + k] = v;
        }
    }
}

}  // namespace curc
