/*
 * CPU oracle for the CURC pseudogenome hot path -- TEST INFRASTRUCTURE ONLY
 * (see oracle.h).  Every function cites the reference lines it restates.
 */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#define ORC_PARALLEL_FOR _Pragma("omp parallel for schedule(dynamic, 1024)")
#else
#define ORC_PARALLEL_FOR
#endif

// Stable sort: libstdc++'s parallel multiway mergesort in the OpenMP build
// (a stable sort has one result, so both builds agree byte for byte).
template <class It, class Cmp>
void orc_stable_sort(It a, It b, Cmp cmp) {
#ifdef _OPENMP
    __gnu_parallel::stable_sort(a, b, cmp);
#else
    std::stable_sort(a, b, cmp);
#endif
}

namespace {

constexpr uint32_t INV = 0xFFFFFFFFu;      // constant.hpp:48
constexpr uint16_t INVOFF = 0xFFFFu;       // constant.hpp:49
constexpr uint32_t MAP_K = 28;             // Param.hpp:73 k1
constexpr uint32_t MAP_REF_STEP = 5;       // Param.hpp:73 ref_index_step1
constexpr uint32_t MAP_READ_STEP = 2;      // Param.hpp:73 read_index_step1
constexpr uint32_t MAP_TARGET = 1;         // Param.hpp:73 target_mismatch_count1
constexpr uint32_t MAP_BUCKET = 12;        // Param.hpp:72 bucket_limit

// util.cuh:93-110 / 148-165: A/a=0 C/c=1 G/g=2 T/t=3, anything else 4 (masked to 0).
inline uint8_t enc(char c) {
    switch (c) {
        case 'A': case 'a': return 0;
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'T': case 't': return 3;
        default: return 4;
    }
}
const char DECODE[] = "ACGT";  // util.cuh:92 (first four symbols)

// util.cuh:82-89 (MurmurHash3 fmix64)
inline uint64_t murmur(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

// util.cuh:72-80
inline uint32_t hash_size(uint64_t n) {
    uint32_t hs;
    uint32_t i = 24;
    do {
        hs = 1u << (i++);
    } while (i <= 31 && hs < n);
    return hs;
}

// util.cuh:111-119 (extract_word_cpu): n <= 32 bases starting at base `off`
// of the packed read beginning at word `start`, right-aligned.
inline uint64_t xw(const uint64_t *bits, uint64_t start, uint64_t off, uint32_t n) {
    uint64_t o1 = off >> 5;
    uint32_t o2 = (uint32_t)((off & 31) << 1);
    uint64_t left = bits[start + o1] << o2;
    if ((64u - o2) >= (n << 1)) return left >> ((32u - n) << 1);
    uint64_t right = (bits[start + o1 + 1] >> (62u - o2)) >> 2;
    return (left | right) >> ((32u - n) << 1);
}

// util.cuh:120-127 (extract_word_unaligned): always reads the following word.
inline uint64_t xwu(const uint64_t *bits, uint64_t base_off, uint32_t n) {
    uint64_t o1 = base_off >> 5;
    uint32_t o2 = (uint32_t)((base_off & 31) << 1);
    uint64_t left = bits[o1] << o2;
    uint64_t right = (bits[o1 + 1] >> (62u - o2)) >> 2;
    return (left | right) >> ((32u - n) << 1);
}

// util.cuh:128-130
inline char base_at(const uint64_t *bits, uint64_t start, uint64_t off) {
    return DECODE[(bits[start + (off >> 5)] >> (((~off) & 31) << 1)) & 3];
}
// util.cuh:196-200 (extract_base_unaligned_gpu)
inline char base_at_u(const uint64_t *bits, uint64_t off) {
    return DECODE[(bits[off >> 5] >> (((~off) & 31) << 1)) & 3];
}

// util.cuh:131-144
inline char rev_base(char b) {
    switch (b) {
        case 'a': return 't'; case 'A': return 'T';
        case 'c': return 'g'; case 'C': return 'G';
        case 'g': return 'c'; case 'G': return 'C';
        case 't': return 'a'; case 'T': return 'A';
        default: return 'N';
    }
}

inline uint64_t bswap64(uint64_t v) { return __builtin_bswap64(v); }

// util.cuh:220-226
inline uint64_t revcomp(uint64_t kmer, uint32_t k) {
    kmer = ~kmer;
    kmer = ((kmer & 0x3333333333333333ULL) << 2) | ((kmer & 0xCCCCCCCCCCCCCCCCULL) >> 2);
    kmer = ((kmer & 0x0F0F0F0F0F0F0F0FULL) << 4) | ((kmer & 0xF0F0F0F0F0F0F0F0ULL) >> 4);
    kmer = bswap64(kmer);
    return kmer >> ((32u - k) << 1);
}

// util.cuh:228-232
inline int mm_words(uint64_t a, uint64_t b) {
    uint64_t x = a ^ b;
    uint64_t r = ((x & 0xAAAAAAAAAAAAAAAAULL) >> 1) | (x & 0x5555555555555555ULL);
    return __builtin_popcountll(r);
}

// util.cuh:233-250; data2 is the packed read (start word 0).
inline int fwd_mm(const uint64_t *ref, uint64_t o1, const uint64_t *rd, uint32_t L, uint32_t bound) {
    int mm = 0;
    uint32_t i = 0;
    for (i = 0; (i + 32 < L) && (mm <= (int)bound); i += 32)
        mm += mm_words(xwu(ref, o1 + i, 32), xw(rd, 0, i, 32));
    if ((i < L) && (mm <= (int)bound))
        mm += mm_words(xwu(ref, o1 + i, L - i), xw(rd, 0, i, L - i));
    return mm;
}

// util.cuh:251-270
inline int rc_mm(const uint64_t *ref, uint64_t o1, const uint64_t *rd, uint32_t L, uint32_t bound) {
    int mm = 0;
    uint32_t i = 0;
    for (i = 0; (i + 32 < L) && (mm <= (int)bound); i += 32)
        mm += mm_words(xwu(ref, o1 + i, 32), revcomp(xw(rd, 0, L - (i + 32), 32), 32));
    if ((i < L) && (mm <= (int)bound))
        mm += mm_words(xwu(ref, o1 + i, L - i), revcomp(xw(rd, 0, 0, L - i), L - i));
    return mm;
}

// util.cuh:271-281 (get_kmer_integer): ASCII k-mer, non-ACGT packs as 0.
inline uint64_t kmer_ascii(const char *s, uint32_t k) {
    uint64_t bits = 0;
    uint32_t idx2 = 0;
    uint32_t kb = k * 2;
    for (uint32_t i = 0; i < k; ++i) {
        uint64_t c = enc(s[i]) & 3u;
        bits |= c << (kb - 2 - idx2);
        idx2 = (idx2 + 2) & 63u;
    }
    return bits;
}

// compress.cu:2159-2174 (encode): MSB-first 2-bit pack, tail bits zero.
void pack_reads(const char *ascii, uint64_t n, uint32_t L, uint32_t W, uint64_t *out) {
    ORC_PARALLEL_FOR
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        uint64_t *dst = out + (uint64_t)i * W;
        for (uint32_t w = 0; w < W; ++w) dst[w] = 0;
        const char *src = ascii + (uint64_t)i * L;
        uint32_t idx1 = 0, idx2 = 0;
        for (uint32_t j = 0; j < L; ++j) {
            uint64_t c = enc(src[j]) & 3u;
            dst[idx1] |= c << (62u - idx2);
            idx2 = (idx2 + 2) & 63u;
            if (idx2 == 0) idx1++;
        }
    }
}

inline uint64_t rec_word(uint64_t isN, uint64_t strand, uint64_t pos) {
    return (isN & 1) | ((strand & 1) << 1) | (pos << 2);
}
inline uint64_t rec_pos(const orc_rec &r) { return r.word >> 2; }

struct PG {
    std::string text;
    std::vector<orc_rec> rec;
};

void push_rec(PG &pg, uint32_t id, uint64_t isN) {
    orc_rec r;
    r.read_id = id;
    r.pad = 0;
    r.word = rec_word(isN, 0, pg.text.size());
    pg.rec.push_back(r);
}

// compress.cu:107-165 (RefRecord::assembly over packed reads)
void assemble_packed(PG &pg, const uint64_t *reads, uint32_t W, uint32_t L,
                     const uint32_t *visit_ids, uint64_t n_visit,
                     const uint32_t *next, const uint32_t *prev, const uint16_t *offset,
                     uint64_t n_reads) {
    std::vector<char> visit(n_reads, 0);
    auto emit = [&](uint32_t id, uint32_t nb) {
        push_rec(pg, id, 0);
        for (uint32_t j = 0; j < nb; ++j) pg.text.push_back(base_at(reads, (uint64_t)id * W, j));
    };
    for (uint64_t r = 0; r < n_visit; ++r) {
        uint32_t rid = visit_ids[r];
        if (visit[rid]) continue;
        uint32_t i = rid;
        while (prev[i] != INV && prev[i] != rid) i = prev[i];
        if (prev[i] == rid) {
            uint16_t mo = offset[i];
            uint32_t mo_id = i;
            while (i != rid) {
                visit[i] = 1;
                i = next[i];
                if (offset[i] > mo) {
                    mo = offset[i];
                    mo_id = i;
                }
            }
            visit[rid] = 1;
            uint32_t s = next[mo_id];
            while (s != mo_id) {
                emit(s, offset[s]);
                s = next[s];
            }
            emit(mo_id, L);
        } else {
            while (next[i] != INV) {
                visit[i] = 1;
                emit(i, offset[i]);
                i = next[i];
            }
            visit[i] = 1;
            emit(i, L);
        }
    }
}

// compress.cu:167-217 (RefRecord::assembly over N-read ASCII)
void assemble_ascii(PG &pg, const char *nreads, uint32_t L, const uint32_t *visit_ids,
                    uint64_t n_visit, const uint32_t *next, const uint32_t *prev,
                    const uint16_t *offset, uint64_t n_reads) {
    std::vector<char> visit(n_reads, 0);
    auto emit = [&](uint32_t id, uint32_t nb) {
        push_rec(pg, id, 1);
        pg.text.append(nreads + (uint64_t)id * L, nb);
    };
    for (uint64_t r = 0; r < n_visit; ++r) {
        uint32_t rid = visit_ids[r];
        if (visit[rid]) continue;
        uint32_t i = rid;
        while (prev[i] != INV && prev[i] != rid) i = prev[i];
        if (prev[i] == rid) {
            uint16_t mo = offset[i];
            uint32_t mo_id = i;
            while (i != rid) {
                visit[i] = 1;
                i = next[i];
                if (offset[i] > mo) {
                    mo = offset[i];
                    mo_id = i;
                }
            }
            visit[rid] = 1;
            uint32_t s = next[mo_id];
            while (s != mo_id) {
                emit(s, offset[s]);
                s = next[s];
            }
            emit(mo_id, L);
        } else {
            while (next[i] != INV) {
                visit[i] = 1;
                emit(i, offset[i]);
                i = next[i];
            }
            visit[i] = 1;
            emit(i, L);
        }
    }
}

// compress.cu:229-246 (compute_binary_ref): ceil(len/32)+1 words, 2-bit MSB-first.
std::vector<uint64_t> pack_pg(const std::string &s) {
    std::vector<uint64_t> out((s.size() + 31) / 32 + 1, 0);
    for (size_t i = 0; i < s.size(); ++i) {
        uint64_t c = enc(s[i]) & 3u;
        out[i / 32] |= c << (62u - 2u * (uint32_t)(i % 32));
    }
    return out;
}

// Lexicographic compare of a[ao : ao+len) vs b[bo : bo+len) on packed reads,
// 32 bases at a time (compress.cu:1484-1493, 1027-1035).
inline int cmp_bases(const uint64_t *reads, uint32_t W, uint32_t a, uint32_t ao, uint32_t b,
                     uint32_t bo, uint32_t len) {
    for (uint32_t j = 0; j < len; j += 32) {
        uint32_t n = std::min<uint32_t>(32, len - j);
        uint64_t x = xw(reads, (uint64_t)a * W, ao + j, n);
        uint64_t y = xw(reads, (uint64_t)b * W, bo + j, n);
        if (x < y) return -1;
        if (x > y) return 1;
    }
    return 0;
}

// compress.cu:1003-1174 (full_match_gpu), restated sequentially.  `pre` and
// `suf` arrive in dedup (lexicographic) order and have equal size (every
// link adds one next and one prev, so heads and tails stay equinumerous).
void full_match_packed(const uint64_t *reads, uint32_t W, uint32_t L, std::vector<uint32_t> pre,
                       std::vector<uint32_t> suf, uint32_t *next, uint32_t *prev,
                       uint16_t *offset) {
    std::vector<uint32_t> tmp1, tmp2;
    std::vector<uint32_t> pr_range, su_range;
    for (uint32_t off = 1; off < L; ++off) {
        // compress.cu:1015-1025: parts by base at off-1 (list is sorted by it).
        auto part_begin = [&](char c) {
            return (size_t)(std::lower_bound(suf.begin(), suf.end(), c,
                                             [&](uint32_t id, char v) {
                                                 return base_at(reads, (uint64_t)id * W, off - 1) < v;
                                             }) -
                            suf.begin());
        };
        size_t A = part_begin('A'), C = part_begin('C'), G = part_begin('G'), T = part_begin('T');
        (void)A;
        // compress.cu:1026-1059: three stable merges by read[off:L].
        auto less = [&](uint32_t a, uint32_t b) {
            return cmp_bases(reads, W, a, off, b, off, L - off) < 0;
        };
        tmp1.resize(G);
        tmp2.resize(suf.size() - G);
        std::merge(suf.begin(), suf.begin() + C, suf.begin() + C, suf.begin() + G, tmp1.begin(), less);
        std::merge(suf.begin() + G, suf.begin() + T, suf.begin() + T, suf.end(), tmp2.begin(), less);
        std::merge(tmp1.begin(), tmp1.end(), tmp2.begin(), tmp2.end(), suf.begin(), less);

        // compress.cu:1061-1088: blocks by the first bp bases.  lower_bound per
        // block id is computed here by one linear pass over each sorted list.
        uint32_t bp = (L - off >= 10) ? 10 : (L - off);
        size_t nb = (size_t)1 << (2 * bp);
        auto ranges = [&](const std::vector<uint32_t> &lst, uint32_t at, std::vector<uint32_t> &rg) {
            rg.assign(nb + 1, 0);
            size_t b = 0;
            for (size_t i = 0; i < lst.size(); ++i) {
                uint64_t id = xw(reads, (uint64_t)lst[i] * W, at, bp);
                while (b <= id) rg[b++] = (uint32_t)i;
            }
            while (b <= nb) rg[b++] = (uint32_t)lst.size();
        };
        ranges(pre, 0, pr_range);
        ranges(suf, off, su_range);

        // compress.cu:1089-1160: per-block merge join with self-avoid rotation.
        uint32_t len = L - off;
        auto cmp_tail = [&](uint32_t p, uint32_t s) {
            // compares p[bp:len) with s[off+bp : off+len)
            return cmp_bases(reads, W, p, bp, s, off + bp, len - bp);
        };
        ORC_PARALLEL_FOR
        for (int64_t bi = 0; bi < (int64_t)nb; ++bi) {
            uint32_t pb = pr_range[bi], pe = pr_range[bi + 1];
            uint32_t sb = su_range[bi], se = su_range[bi + 1];
            uint32_t pi = pb, si = sb;
            while (pi < pe && si < se) {
                uint32_t pid = pre[pi];
                uint32_t sid = suf[si];
                int c = cmp_tail(pid, sid);
                if (c > 0) {
                    si++;
                } else if (c < 0) {
                    pi++;
                } else {
                    if (pid == sid) {
                        uint32_t cur = pi;
                        int t = 1;
                        while (pi + 1 < pe) {
                            pi++;
                            pid = pre[pi];
                            // compress.cu:1119-1128 only tests '>' (list is sorted).
                            t = cmp_tail(pid, sid) > 0 ? 1 : 0;
                            if (t != 0) break;
                            if (pid != sid) break;
                            t = 1;
                        }
                        if (t != 0) {
                            pi = cur;
                            si++;
                            continue;
                        } else {
                            pid = pre[pi];
                            while (pi > cur) {
                                pre[pi] = pre[pi - 1];
                                pi--;
                            }
                            pre[pi] = pid;
                        }
                    }
                    next[sid] = pid;
                    prev[pid] = sid;
                    pi++;
                    si++;
                }
            }
        }
        // compress.cu:1162-1172
        for (uint32_t id : suf)
            if (next[id] != INV) offset[id] = (uint16_t)off;
        pre.erase(std::remove_if(pre.begin(), pre.end(), [&](uint32_t id) { return prev[id] != INV; }),
                  pre.end());
        suf.erase(std::remove_if(suf.begin(), suf.end(), [&](uint32_t id) { return next[id] != INV; }),
                  suf.end());
    }
}

// compress.cu:1176-1347 (full_match_cpu) over N-read ASCII, alphabet ACGNT.
// `ids` (ascending N indices) is sorted in place; ties by ascending index.
void full_match_ascii(const char *nr, uint32_t L, std::vector<uint32_t> &ids, uint64_t n_N,
                      std::vector<uint32_t> &next, std::vector<uint32_t> &prev,
                      std::vector<uint16_t> &offset) {
    auto rd = [&](uint32_t id) { return nr + (uint64_t)id * L; };
    std::stable_sort(ids.begin(), ids.end(),
                     [&](uint32_t a, uint32_t b) { return std::memcmp(rd(a), rd(b), L) < 0; });
    prev.assign(n_N, INV);
    next.assign(n_N, INV);
    offset.assign(n_N, INVOFF);
    for (size_t i = 1; i < ids.size(); ++i) {
        uint32_t a = ids[i - 1], b = ids[i];
        if (std::memcmp(rd(a), rd(b), L) == 0) {
            prev[b] = a;
            next[a] = b;
            offset[a] = 0;
        }
    }
    std::vector<uint32_t> pre, last_suf, suf;
    for (uint32_t id : ids) {
        if (prev[id] == INV) pre.push_back(id);
        if (next[id] == INV) last_suf.push_back(id);
    }
    suf.resize(last_suf.size());
    static const char ALPHA[5] = {'A', 'C', 'G', 'N', 'T'};
    auto digit = [](char c) -> int {
        switch (c) {
            case 'A': return 0; case 'C': return 1; case 'G': return 2;
            case 'N': return 3; case 'T': return 4;
            default: return -1;
        }
    };
    (void)ALPHA;
    uint32_t off = 1;
    while (true) {
        // compress.cu:1217-1242: 5 parts by base at off-1, stable multiway merge.
        size_t cnt[5] = {0, 0, 0, 0, 0};
        for (uint32_t id : last_suf) cnt[digit(rd(id)[off - 1])]++;
        size_t bnd[6];
        bnd[0] = 0;
        for (int c = 0; c < 5; ++c) bnd[c + 1] = bnd[c] + cnt[c];
        auto less = [&](uint32_t a, uint32_t b) {
            return std::memcmp(rd(a) + off, rd(b) + off, L - off) < 0;
        };
        // stable k-way merge: repeatedly take the smallest head; ties go to the
        // lower part index (stable_multiway_merge semantics).
        {
            size_t h[5];
            for (int c = 0; c < 5; ++c) h[c] = bnd[c];
            for (size_t o = 0; o < suf.size(); ++o) {
                int best = -1;
                for (int c = 0; c < 5; ++c) {
                    if (h[c] == bnd[c + 1]) continue;
                    if (best < 0 || less(last_suf[h[c]], last_suf[h[best]])) best = c;
                }
                suf[o] = last_suf[h[best]++];
            }
        }
        // compress.cu:1244-1271: blocks by the first bp (<=3) chars, base 5.
        uint32_t bp = (L - off >= 3) ? 3 : (L - off);
        size_t nb = 1;
        for (uint32_t i = 0; i < bp; ++i) nb *= 5;
        auto block_id = [&](const char *s) {
            size_t v = 0;
            for (uint32_t i = 0; i < bp; ++i) v = v * 5 + (size_t)digit(s[i]);
            return v;
        };
        std::vector<uint32_t> pr(nb + 1), sr(nb + 1);
        auto ranges = [&](const std::vector<uint32_t> &lst, uint32_t at, std::vector<uint32_t> &rg) {
            size_t b = 0;
            for (size_t i = 0; i < lst.size(); ++i) {
                size_t id = block_id(rd(lst[i]) + at);
                while (b <= id) rg[b++] = (uint32_t)i;
            }
            while (b <= nb) rg[b++] = (uint32_t)lst.size();
        };
        ranges(pre, 0, pr);
        ranges(suf, off, sr);
        // compress.cu:1273-1325
        uint32_t tl = L - off - bp;
        for (size_t bi = 0; bi < nb; ++bi) {
            uint32_t pi = pr[bi], pe = pr[bi + 1], si = sr[bi], se = sr[bi + 1];
            while (pi < pe && si < se) {
                uint32_t pid = pre[pi], sid = suf[si];
                int c = std::memcmp(rd(pid) + bp, rd(sid) + off + bp, tl);
                if (c > 0) {
                    si++;
                } else if (c < 0) {
                    pi++;
                } else {
                    if (pid == sid) {
                        uint32_t cur = pi;
                        int t = 1;
                        while (pi + 1 < pe) {
                            pi++;
                            pid = pre[pi];
                            t = std::memcmp(rd(pid) + bp, rd(sid) + off + bp, tl);
                            if (t != 0) break;
                            if (pid != sid) break;
                            t = 1;
                        }
                        if (t != 0) {
                            pi = cur;
                            si++;
                            continue;
                        } else {
                            pid = pre[pi];
                            while (pi > cur) {
                                pre[pi] = pre[pi - 1];
                                pi--;
                            }
                            pre[pi] = pid;
                        }
                    }
                    next[sid] = pid;
                    prev[pid] = sid;
                    pi++;
                    si++;
                }
            }
        }
        for (uint32_t id : suf)
            if (next[id] != INV) offset[id] = (uint16_t)off;
        pre.erase(std::remove_if(pre.begin(), pre.end(), [&](uint32_t id) { return prev[id] != INV; }),
                  pre.end());
        suf.erase(std::remove_if(suf.begin(), suf.end(), [&](uint32_t id) { return next[id] != INV; }),
                  suf.end());
        last_suf = suf;
        off++;
        if (off >= L) break;
    }
}

struct MapIndex {
    uint32_t H = 0;
    std::vector<uint32_t> kr;   // H + 2
    std::vector<uint32_t> pos;  // bucket contents (index i, position = start + i*step)
};

// compress.cu:362-409 (build_index), sequential meaning of the racy uint8
// counters: the first min(n, 12) occurrences by ascending i, stored in
// descending i.
void build_map_index(MapIndex &ix, const uint64_t *ref, uint64_t start, uint64_t count, bool rc) {
    ix.H = hash_size(count);
    uint32_t mask = ix.H - 1;
    std::vector<uint8_t> cnt((size_t)ix.H + 2, 0);
    std::vector<uint32_t> hv(count);
    ORC_PARALLEL_FOR
    for (int64_t i = 0; i < (int64_t)count; ++i) {
        uint64_t km = xwu(ref, start + (uint64_t)i * MAP_REF_STEP, MAP_K);
        if (rc) km = revcomp(km, MAP_K);
        hv[i] = (uint32_t)(murmur(km) & mask);
    }
    for (uint64_t i = 0; i < count; ++i)
        if (cnt[hv[i]] < MAP_BUCKET) ++cnt[hv[i]];
    ix.kr.assign((size_t)ix.H + 2, 0);
    uint32_t acc = 0;
    for (size_t h = 0; h < (size_t)ix.H + 2; ++h) {
        ix.kr[h] = acc;
        acc += cnt[h];
    }
    ix.pos.assign((size_t)acc + 2, 0);
    for (uint64_t i = 0; i < count; ++i) {
        uint32_t h = hv[i];
        if (cnt[h]) ix.pos[ix.kr[h] + (--cnt[h])] = (uint32_t)i;
    }
}

// Not part of the reference: the early end of a strand pass that the
// library's mapping kernel uses (a pass stops once every class of probes,
// t mod 5, has had more than 3 (last - 1) probes of buckets holding fewer
// than MAP_BUCKET positions; see k_map_both).  Off by default; the tests turn
// it on to show that it never changes a result of the loop below.
// g_map_settle_kill: the probes of one class a mismatch is taken to spoil
// (3 is the rule; 0 turns the early end off; smaller values are mutants).
uint32_t g_map_settle_kill = 0;
std::atomic<uint64_t> g_map_settled{0};  // passes ended early

bool map_settled(const uint32_t clean[5], uint32_t last) {
    if (last <= 1) return true;
    const uint32_t need = g_map_settle_kill * (last - 1) + 1;
    for (int c = 0; c < 5; ++c)
        if (clean[c] < need) return false;
    return true;
}

// compress.cu:411-521 (mapping<strand>) for one index half.  settle: apply
// the early end above (one index half only).
void map_reads(const MapIndex &ix, uint64_t start, bool rc, const uint64_t *ref, uint64_t ref_len,
               uint32_t L, uint32_t W, const uint64_t *gathered, uint64_t n_un, const char *nascii,
               uint64_t n_N, std::vector<uint8_t> &last, std::vector<orc_rec> &m, bool settle = false) {
    uint32_t mask = ix.H - 1;
    ORC_PARALLEL_FOR
    for (int64_t ii = 0; ii < (int64_t)(n_un + n_N); ++ii) {
        uint64_t i = (uint64_t)ii;
        if (last[i] <= MAP_TARGET) continue;
        bool isN = i >= n_un;
        const uint64_t *rd = isN ? nullptr : gathered + i * W;
        const char *ra = isN ? nascii + (i - n_un) * L : nullptr;
        uint32_t clean[5] = {0, 0, 0, 0, 0};
        for (uint32_t j = 0; j + MAP_K <= L; j += MAP_READ_STEP) {
            if (settle && j / MAP_READ_STEP >= 5 * (g_map_settle_kill + 1) && map_settled(clean, last[i])) {
                g_map_settled.fetch_add(1, std::memory_order_relaxed);
                break;
            }
            uint64_t key = isN ? kmer_ascii(ra + j, MAP_K) : xw(rd, 0, j, MAP_K);
            uint32_t h = (uint32_t)(murmur(key) & mask);
            uint32_t b = ix.kr[h], e = ix.kr[h + 1];
            if (e - b < MAP_BUCKET) clean[(j / MAP_READ_STEP) % 5]++;
            if (b == e) continue;
            e = std::min(b + MAP_BUCKET, e);
            for (uint32_t p = b; p < e; ++p) {
                uint64_t rpos = start + (uint64_t)ix.pos[p] * MAP_REF_STEP;
                uint64_t ls = rc ? (uint64_t)(L - MAP_K - j) : (uint64_t)j;
                if (rpos < ls) continue;
                uint64_t mp = rpos - ls;
                if (mp + L > ref_len) continue;
                uint32_t cnt;
                if (!isN) {
                    cnt = (uint32_t)(rc ? rc_mm(ref, mp, rd, L, last[i] - 1)
                                        : fwd_mm(ref, mp, rd, L, last[i] - 1));
                } else {
                    cnt = 0;
                    for (uint32_t o = 0; o < L; ++o) {
                        char rb = base_at_u(ref, mp + o);
                        char qb = rc ? rev_base(ra[L - 1 - o]) : ra[o];
                        if (rb != qb) {
                            cnt++;
                            if (cnt >= last[i]) break;
                        }
                    }
                }
                if (cnt < last[i]) {
                    m[i].word = rec_word(isN ? 1 : 0, rc ? 1 : 0, mp);
                    last[i] = (uint8_t)cnt;
                    if (last[i] <= MAP_TARGET) goto done;
                }
            }
        }
    done:;
    }
}

// Not part of the reference: the two strand passes of a read interleaved
// probe by probe, as the library's k_map_both runs them (the reverse pass
// accepts a count below min(its best, the forward best so far) and wins when
// below the final forward best; a forward pass reaching MAP_TARGET ends the
// read), with the early end of each pass when g_map_settle_kill > 0: the
// forward pass with B = min(forward best, max(reverse best + 1,
// MAP_TARGET + 1)) (a forward window that does not reach the reverse best
// cannot change the answer, unless it reaches MAP_TARGET: then the forward
// pass ends the read and the reverse pass never counts), the reverse pass
// with B = min(reverse best, forward best).  The tests compare it with
// map_reads over both strands.  g_map_both_tie: a mutant that drops the + 1.
bool g_map_both = false;
bool g_map_both_tie = false;

void map_reads_both(const MapIndex &ixf, const MapIndex &ixr, const uint64_t *ref, uint64_t ref_len, uint32_t L,
                    uint32_t W, const uint64_t *gathered, uint64_t n_un, const char *nascii, uint64_t n_N,
                    std::vector<uint8_t> &last, std::vector<orc_rec> &m) {
    const uint32_t np = (L - MAP_K) / MAP_READ_STEP + 1;
    const uint32_t kill = g_map_settle_kill;
    ORC_PARALLEL_FOR
    for (int64_t ii = 0; ii < (int64_t)(n_un + n_N); ++ii) {
        const uint64_t i = (uint64_t)ii;
        if (last[i] <= MAP_TARGET) continue;
        const bool isN = i >= n_un;
        const uint64_t *rd = isN ? nullptr : gathered + i * W;
        const char *ra = isN ? nascii + (i - n_un) * L : nullptr;
        auto count = [&](bool rc, uint64_t mp, uint32_t bound) -> uint32_t {
            if (!isN) return (uint32_t)(rc ? rc_mm(ref, mp, rd, L, bound) : fwd_mm(ref, mp, rd, L, bound));
            uint32_t c = 0;
            for (uint32_t o = 0; o < L; ++o) {
                char rb = base_at_u(ref, mp + o);
                char qb = rc ? rev_base(ra[L - 1 - o]) : ra[o];
                if (rb != qb && ++c > bound) break;
            }
            return c;
        };
        uint32_t F = last[i], Rb = last[i];
        uint64_t wF = m[i].word, wR = m[i].word;
        bool doneF = false, doneR = false, fwdT = false;
        uint32_t cleanF[5] = {0, 0, 0, 0, 0}, cleanR[5] = {0, 0, 0, 0, 0};
        for (uint32_t t = 0; t < np && !(doneF && doneR); ++t) {
            const uint32_t j = MAP_READ_STEP * t;
            const uint64_t key = isN ? kmer_ascii(ra + j, MAP_K) : xw(rd, 0, j, MAP_K);
            const uint32_t hv = (uint32_t)murmur(key);
            if (!doneF) {
                const uint32_t h = hv & (ixf.H - 1), b = ixf.kr[h], e = std::min(b + MAP_BUCKET, ixf.kr[h + 1]);
                if (ixf.kr[h + 1] - b < MAP_BUCKET) cleanF[t % 5]++;
                for (uint32_t p = b; p < e && !fwdT; ++p) {
                    const uint64_t rpos = (uint64_t)ixf.pos[p] * MAP_REF_STEP;
                    if (rpos < j) continue;
                    const uint64_t mp = rpos - j;
                    if (mp + L > ref_len) continue;
                    const uint32_t c = count(false, mp, F - 1);
                    if (c < F) {
                        F = c;
                        wF = rec_word(isN ? 1 : 0, 0, mp);
                        if (F <= MAP_TARGET) fwdT = true;
                    }
                }
                if (fwdT) break;
            }
            if (!doneR) {
                const uint32_t h = hv & (ixr.H - 1), b = ixr.kr[h], e = std::min(b + MAP_BUCKET, ixr.kr[h + 1]);
                if (ixr.kr[h + 1] - b < MAP_BUCKET) cleanR[t % 5]++;
                const uint64_t ls = L - MAP_K - j;
                for (uint32_t p = b; p < e; ++p) {
                    const uint64_t rpos = (uint64_t)ixr.pos[p] * MAP_REF_STEP;
                    if (rpos < ls) continue;
                    const uint64_t mp = rpos - ls;
                    if (mp + L > ref_len) continue;
                    const uint32_t lim = std::min(Rb, F);
                    const uint32_t c = count(true, mp, lim - 1);
                    if (c < lim) {
                        Rb = c;
                        wR = rec_word(isN ? 1 : 0, 1, mp);
                        if (Rb <= MAP_TARGET) {
                            doneR = true;
                            break;
                        }
                    }
                }
            }
            if (kill) {
                const uint32_t bf = std::min(F, std::max(g_map_both_tie ? Rb : Rb + 1, MAP_TARGET + 1));
                const uint32_t br = std::min(Rb, F);
                const bool late = t + 1 >= 5 * (kill + 1);
                if (!doneF && (bf <= 1 || (late && map_settled(cleanF, bf)))) {
                    doneF = true;
                    g_map_settled.fetch_add(1, std::memory_order_relaxed);
                }
                if (!doneR && (br <= 1 || (late && map_settled(cleanR, br)))) {
                    doneR = true;
                    g_map_settled.fetch_add(1, std::memory_order_relaxed);
                }
            }
        }
        if (!fwdT && Rb < F) {
            last[i] = (uint8_t)Rb;
            m[i].word = wR;
        } else {
            last[i] = (uint8_t)F;
            m[i].word = wF;
        }
    }
}

}  // namespace

struct orc_block {
    orc_params P{};
    uint64_t n_reads = 0;
    std::vector<uint64_t> packed;
    std::vector<uint32_t> nonN, Nids, sorted;
    std::vector<char> Nascii;
    std::vector<uint32_t> next_dedup, prev_dedup, next_ov, prev_ov, next_pr, prev_pr;
    std::vector<uint16_t> off_ov;
    std::vector<uint32_t> next, prev;
    std::vector<uint16_t> offset;
    std::vector<uint32_t> hq, unmatched;
    PG hq_pg, lq_pg, nlq_pg;
    std::vector<uint64_t> hq_pg2;
    std::vector<orc_rec> map_rec;
    std::vector<uint8_t> map_mis;
    std::vector<uint32_t> unmapped, unmapped_N;
    std::vector<uint32_t> lq_next, lq_prev;
    std::vector<uint16_t> lq_off;
    // pre-entropy streams (orc_block_streams)
    std::string s_strand;
    std::vector<uint16_t> s_read_off, s_lq_off, s_nlq_off;
    std::vector<uint8_t> s_mis_count, s_mis_base, s_mis_table;
    std::vector<uint32_t> s_pe_order;
    std::vector<uint64_t> s_id_pos;
    std::vector<orc_rec> s_sorted;
    std::vector<std::vector<uint16_t>> s_mis_off;
};

extern "C" {

// The early end of strand passes in the mapping (see map_settled) for the
// blocks run after this call: kill = probes of a class one mismatch spoils
// (3: the library's rule; 0: off, the reference's exhaustive loop).
void orc_set_map_settle(int kill) {
    g_map_settle_kill = kill > 0 ? (uint32_t)kill : 0u;
    g_map_settled.store(0);
}
// Strand passes ended early since the last orc_set_map_settle.
uint64_t orc_map_settled_count() { return g_map_settled.load(); }
// Both strand passes interleaved (map_reads_both) instead of one after the
// other; tie != 0 is the mutant without the + 1 of the forward bound.
void orc_set_map_both(int on, int tie) {
    g_map_both = on != 0;
    g_map_both_tie = tie != 0;
}

int orc_make_params(uint32_t L, orc_params *p) {
    // preprocess.cpp:296-313
    if (L == 0 || L > 512) return -1;
    p->L = L;
    p->W = (L + 31) / 32;
    p->max_off = L * 65 / 100;
    p->k = 32;
    if (L < 89) p->k = 36 * L / 100;
    if (p->k == 0 || p->k + p->max_off > L) return -1;
    p->kip = L - p->k - p->max_off;
    p->max_mm = L / 5;
    return 0;
}

uint64_t orc_murmur64(uint64_t x) { return murmur(x); }
uint32_t orc_hash_size(uint64_t n) { return hash_size(n); }
uint64_t orc_extract_word(const uint64_t *bits, uint64_t start, uint64_t off, uint32_t n) {
    return xw(bits, start, off, n);
}
uint64_t orc_extract_word_unaligned(const uint64_t *bits, uint64_t base_off, uint32_t n) {
    return xwu(bits, base_off, n);
}
uint64_t orc_revcomp_kmer(uint64_t kmer, uint32_t k) { return revcomp(kmer, k); }
int orc_mismatch_words(uint64_t a, uint64_t b) { return mm_words(a, b); }
int orc_forward_mismatch(const uint64_t *ref, uint64_t ref_off, const uint64_t *read, uint32_t L,
                         uint32_t bound) {
    return fwd_mm(ref, ref_off, read, L, bound);
}
int orc_rc_mismatch(const uint64_t *ref, uint64_t ref_off, const uint64_t *read, uint32_t L,
                    uint32_t bound) {
    return rc_mm(ref, ref_off, read, L, bound);
}
void orc_pack(const char *ascii, uint64_t n, uint32_t L, uint64_t *packed) {
    pack_reads(ascii, n, L, (L + 31) / 32, packed);
}

// The record loop of process() (preprocess.cpp:147-193) over one block's
// bytes: per record, skip the header line, take the sequence line (its
// length must be L, :152-154), skip the '+' and quality lines; a paired
// block reads file 2's record in step with file 1's, ids 2k / 2k+1
// (:156-192).  The loop runs over file 1's bytes as the reference's does;
// a file 2 that runs out of records first is reported (-3) where the
// reference would read past its buffer.  Reads land in `out` in id order
// (cap_reads of them at most).  Returns 0, -1 (read length isn't fix),
// -2 (truncated record) or -3 (record counts differ) / -4 (cap exceeded).
int orc_parse_fastq(const char *fq1, uint64_t bytes1, const char *fq2, uint64_t bytes2, uint32_t L, char *out,
                    uint64_t cap_reads, uint64_t *n_reads) {
    uint64_t i = 0, j = 0, n = 0;
    const bool paired = fq2 != nullptr;
    auto line_end = [](const char *b, uint64_t len, uint64_t at) -> uint64_t {
        const void *p = at < len ? std::memchr(b + at, '\n', len - at) : nullptr;
        return p ? (uint64_t)((const char *)p - b) : len;
    };
    *n_reads = 0;
    while (i < bytes1) {
        const uint64_t h = line_end(fq1, bytes1, i);
        const uint64_t seq = h + 1;
        const uint64_t e = line_end(fq1, bytes1, seq);
        if (e >= bytes1) return -2;
        if (e - seq != L) return -1;
        if (n >= cap_reads) return -4;
        std::memcpy(out + n * L, fq1 + seq, L);
        ++n;
        const uint64_t plus = line_end(fq1, bytes1, e + 1);
        const uint64_t qual = plus < bytes1 ? line_end(fq1, bytes1, plus + 1) : bytes1;
        if (qual >= bytes1) return -2;
        i = qual + 1;
        if (paired) {
            if (j >= bytes2) return -3;
            const uint64_t h2 = line_end(fq2, bytes2, j);
            const uint64_t seq2 = h2 + 1;
            const uint64_t e2 = line_end(fq2, bytes2, seq2);
            if (e2 >= bytes2) return -2;
            if (e2 - seq2 != L) return -1;
            if (n >= cap_reads) return -4;
            std::memcpy(out + n * L, fq2 + seq2, L);
            ++n;
            const uint64_t plus2 = line_end(fq2, bytes2, e2 + 1);
            const uint64_t qual2 = plus2 < bytes2 ? line_end(fq2, bytes2, plus2 + 1) : bytes2;
            if (qual2 >= bytes2) return -2;
            j = qual2 + 1;
        }
    }
    if (paired && j < bytes2) return -3;
    *n_reads = n;
    return 0;
}

orc_block *orc_block_run(const char *ascii, uint64_t n_reads, uint32_t L, int stop_after) {
    orc_params P;
    if (orc_make_params(L, &P) != 0) return nullptr;
    if (n_reads >= (uint64_t)INV) return nullptr;
    orc_block *B = new orc_block();
    B->P = P;
    B->n_reads = n_reads;
    const uint32_t W = P.W;

    // a1: N split (preprocess.cpp:147-193): a read is an N-read iff it holds 'N'.
    for (uint64_t i = 0; i < n_reads; ++i) {
        const char *r = ascii + i * L;
        if (std::memchr(r, 'N', L)) {
            B->Nids.push_back((uint32_t)i);
            B->Nascii.insert(B->Nascii.end(), r, r + L);
        } else {
            B->nonN.push_back((uint32_t)i);
        }
    }
    // a2: every read (N-reads too) is packed (preprocess.cpp:186-199).
    B->packed.assign(n_reads * W + 1, 0);
    pack_reads(ascii, n_reads, L, W, B->packed.data());
    const uint64_t *R = B->packed.data();

    // a4: dedup sort + links (compress.cu:1396-1420), ties by ascending id.
    B->next.assign(n_reads, INV);
    B->prev.assign(n_reads, INV);
    B->offset.assign(n_reads, INVOFF);
    uint32_t *next = B->next.data(), *prev = B->prev.data();
    uint16_t *offset = B->offset.data();
    B->sorted = B->nonN;
    orc_stable_sort(B->sorted.begin(), B->sorted.end(), [&](uint32_t a, uint32_t b) {
        for (uint32_t w = 0; w < W; ++w) {
            uint64_t x = R[(uint64_t)a * W + w], y = R[(uint64_t)b * W + w];
            if (x != y) return x < y;
        }
        return false;
    });
    const std::vector<uint32_t> &S = B->sorted;
    for (size_t i = 0; i + 1 < S.size(); ++i) {
        if (std::memcmp(R + (uint64_t)S[i] * W, R + (uint64_t)S[i + 1] * W, W * 8) == 0) {
            next[S[i]] = S[i + 1];
            prev[S[i + 1]] = S[i];
            offset[S[i]] = 0;
        }
    }
    B->next_dedup = B->next;
    B->prev_dedup = B->prev;
    if (stop_after <= ORC_STAGE_DEDUP) return B;

    // a5: prefix k-mer index (compress.cu:1427-1471).
    {
        std::vector<uint32_t> value, suffix;
        for (uint32_t id : S) {
            if (prev[id] == INV) value.push_back(id);
            if (next[id] == INV) suffix.push_back(id);
        }
        const uint32_t H = hash_size(value.size());
        const uint32_t mask = H - 1;
        std::vector<uint32_t> hk(value.size());
        for (size_t i = 0; i < value.size(); ++i)
            hk[i] = (uint32_t)(murmur(xw(R, (uint64_t)value[i] * W, P.kip, P.k)) & mask);
        std::vector<uint32_t> ord(value.size());
        std::iota(ord.begin(), ord.end(), 0u);
        orc_stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return hk[a] < hk[b]; });
        std::vector<uint32_t> va(value.size());
        std::vector<uint32_t> kr((size_t)H + 2, 0);
        for (size_t i = 0; i < ord.size(); ++i) {
            va[i] = value[ord[i]];
            kr[hk[ord[i]]]++;
        }
        uint32_t acc = 0;
        for (size_t h = 0; h < (size_t)H + 2; ++h) {
            uint32_t c = kr[h];
            kr[h] = acc;
            acc += c;
        }

        // a6: greedy overlap loop (compress.cu:1473-1532), suffixes in list
        // order, claims visible to later suffixes at once.  The scan of
        // :1502-1527 takes the first candidate of the bucket from the lower
        // bound on that is neither the suffix itself nor taken, and links it
        // if it equals s[off:L].  A bucket is in lexicographic order, so the
        // equal candidates are one run [first, eq) from the lower bound, and
        // that outcome is: the first free non-self member of the run, if any.
        // The run does not depend on the claims, so it is found for all
        // suffixes first (in parallel in the OpenMP build), then the claims
        // are made in list order.
        std::vector<uint32_t> run_lo(suffix.size()), run_hi(suffix.size());
        for (uint32_t off = 1; off < P.max_off; ++off) {
            const uint32_t cl = L - off;
            const int64_t ns = (int64_t)suffix.size();
            ORC_PARALLEL_FOR
            for (int64_t i = 0; i < ns; ++i) {
                const uint32_t s = suffix[i];
                uint64_t km = xw(R, (uint64_t)s * W, P.kip + off, P.k);
                uint32_t h = (uint32_t)(murmur(km) & mask);
                uint32_t first = kr[h], end = kr[h + 1];
                uint32_t count = end - first;
                while (count > 0) {
                    uint32_t step = count / 2;
                    uint32_t it = first + step;
                    if (cmp_bases(R, W, va[it], 0, s, off, cl) < 0) {
                        first = it + 1;
                        count -= step + 1;
                    } else {
                        count = step;
                    }
                }
                uint32_t eq = first;
                while (eq < end && cmp_bases(R, W, va[eq], 0, s, off, cl) == 0) ++eq;
                run_lo[i] = first;
                run_hi[i] = eq;
            }
            for (int64_t i = 0; i < ns; ++i) {
                const uint32_t s = suffix[i];
                for (uint32_t x = run_lo[i]; x < run_hi[i]; ++x) {
                    uint32_t p = va[x];
                    if (p == s || prev[p] != INV) continue;
                    prev[p] = s;
                    next[s] = p;
                    offset[s] = (uint16_t)off;
                    break;
                }
            }
            suffix.erase(std::remove_if(suffix.begin(), suffix.end(),
                                        [&](uint32_t id) { return next[id] != INV; }),
                         suffix.end());
        }
    }
    B->next_ov = B->next;
    B->prev_ov = B->prev;
    B->off_ov = B->offset;
    if (stop_after <= ORC_STAGE_OVERLAP) return B;

    // a7: prune + stable partition (compress.cu:1539-1572).
    {
        std::vector<char> res(n_reads, 1);
        for (uint64_t r = 0; r < n_reads; ++r) {
            if (next[r] != INV && prev[r] != INV) continue;
            if (next[r] != INV && offset[r] == 0) continue;
            if (prev[r] != INV && offset[prev[r]] == 0) continue;
            res[r] = 0;
        }
        for (uint64_t r = 0; r < n_reads; ++r) {
            if (res[r]) continue;
            if (next[r] != INV) {
                prev[next[r]] = INV;
                next[r] = INV;
            }
            if (prev[r] != INV) {
                next[prev[r]] = INV;
                prev[r] = INV;
            }
        }
        for (uint32_t id : S) (res[id] ? B->hq : B->unmatched).push_back(id);
    }
    B->next_pr = B->next;
    B->prev_pr = B->prev;
    if (stop_after <= ORC_STAGE_PRUNE) return B;

    // a8: exact all-offset matching of HQ chain ends (compress.cu:1574-1591).
    {
        std::vector<uint32_t> pre, suf;
        for (uint32_t id : B->hq) {
            if (prev[id] == INV) pre.push_back(id);
            if (next[id] == INV) suf.push_back(id);
        }
        full_match_packed(R, W, L, pre, suf, next, prev, offset);
    }
    if (stop_after <= ORC_STAGE_FULLMATCH) return B;

    // a9: assembly (compress.cu:1613-1618).
    assemble_packed(B->hq_pg, R, W, L, B->hq.data(), B->hq.size(), next, prev, offset, n_reads);
    B->hq_pg2 = pack_pg(B->hq_pg.text);
    if (stop_after <= ORC_STAGE_ASSEMBLY) return B;

    // a10: mapping (compress.cu:1620-1640, 268-595).
    {
        const uint64_t n_un = B->unmatched.size(), n_N = B->Nids.size();
        std::vector<uint64_t> g(n_un * W + 1, 0);
        for (uint64_t i = 0; i < n_un; ++i)
            std::memcpy(&g[i * W], R + (uint64_t)B->unmatched[i] * W, W * 8);
        std::vector<uint8_t> last(n_un + n_N, (uint8_t)(P.max_mm + 1));
        std::vector<orc_rec> m(n_un + n_N);
        for (uint64_t i = 0; i < n_un + n_N; ++i) {
            m[i].read_id = (uint32_t)(i < n_un ? i : i - n_un);
            m[i].pad = 0;
            m[i].word = rec_word(i < n_un ? 0 : 1, 0, 0);
        }
        const uint64_t ref_len = B->hq_pg.text.size();
        const uint64_t *ref = B->hq_pg2.data();
        uint64_t index_count = 0;
        if (ref_len + 1 >= MAP_K) index_count = (ref_len - MAP_K + 1 + MAP_REF_STEP - 1) / MAP_REF_STEP;
        if (L >= MAP_K && g_map_both && index_count <= (1ULL << 30)) {
            // both strands interleaved (test switch, see map_reads_both)
            MapIndex ixf, ixr;
            build_map_index(ixf, ref, 0, index_count, false);
            build_map_index(ixr, ref, 0, index_count, true);
            map_reads_both(ixf, ixr, ref, ref_len, L, W, g.data(), n_un, B->Nascii.data(), n_N, last, m);
        } else if (L >= MAP_K) {
            for (int strand = 0; strand < 2; ++strand) {
                bool rc = strand == 1;
                // compress.cu:341-360: two index halves past 2^30 positions.
                if (index_count > (1ULL << 30)) {
                    uint64_t c1 = index_count / 2, c2 = index_count - c1;
                    uint64_t s2 = c1 * MAP_REF_STEP;
                    MapIndex ix;
                    build_map_index(ix, ref, 0, c1, rc);
                    map_reads(ix, 0, rc, ref, ref_len, L, W, g.data(), n_un, B->Nascii.data(), n_N, last, m);
                    build_map_index(ix, ref, s2, c2, rc);
                    map_reads(ix, s2, rc, ref, ref_len, L, W, g.data(), n_un, B->Nascii.data(), n_N, last, m);
                } else {
                    MapIndex ix;
                    build_map_index(ix, ref, 0, index_count, rc);
                    map_reads(ix, 0, rc, ref, ref_len, L, W, g.data(), n_un, B->Nascii.data(), n_N, last, m,
                              g_map_settle_kill != 0);
                }
            }
        }
        // compress.cu:524-594 (get_result): stable compactions.
        for (uint64_t i = 0; i < n_un + n_N; ++i) {
            bool mapped = last[i] <= P.max_mm;
            if (i < n_un) {
                m[i].read_id = B->unmatched[i];
                if (!mapped) B->unmapped.push_back(B->unmatched[i]);
            } else if (!mapped) {
                B->unmapped_N.push_back((uint32_t)(i - n_un));
            }
        }
        for (uint64_t i = 0; i < n_un + n_N; ++i) {
            if (last[i] <= P.max_mm) {
                B->map_rec.push_back(m[i]);
                B->map_mis.push_back(last[i]);
            }
        }
    }
    if (stop_after <= ORC_STAGE_MAP) return B;

    // a11: LQ pseudogenome from unmapped non-N reads (compress.cu:1649-1722).
    if (!B->unmapped.empty()) {
        const uint64_t n = B->unmapped.size();
        std::vector<uint64_t> g(n * W + 1, 0);
        for (uint64_t i = 0; i < n; ++i) std::memcpy(&g[i * W], R + (uint64_t)B->unmapped[i] * W, W * 8);
        B->lq_next.assign(n, INV);
        B->lq_prev.assign(n, INV);
        B->lq_off.assign(n, INVOFF);
        std::vector<uint32_t> ids(n);
        std::iota(ids.begin(), ids.end(), 0u);
        full_match_packed(g.data(), W, L, ids, ids, B->lq_next.data(), B->lq_prev.data(), B->lq_off.data());
        assemble_packed(B->lq_pg, g.data(), W, L, ids.data(), n, B->lq_next.data(), B->lq_prev.data(),
                        B->lq_off.data(), n);
        for (auto &r : B->lq_pg.rec) r.read_id = B->unmapped[r.read_id];
    }
    if (stop_after <= ORC_STAGE_LQ) return B;

    // a12: N-read LQ pseudogenome (compress.cu:1724-1728, 1176-1347).
    if (!B->unmapped_N.empty()) {
        std::vector<uint32_t> ids = B->unmapped_N;
        std::vector<uint32_t> nx, pv;
        std::vector<uint16_t> of;
        full_match_ascii(B->Nascii.data(), L, ids, B->Nids.size(), nx, pv, of);
        assemble_ascii(B->nlq_pg, B->Nascii.data(), L, ids.data(), ids.size(), nx.data(), pv.data(),
                       of.data(), B->Nids.size());
    }
    return B;
}

uint64_t orc_block_get(orc_block *B, int field, const void **ptr) {
    auto put = [&](const auto &v) -> uint64_t {
        *ptr = v.data();
        return v.size();
    };
    switch (field) {
        case ORC_F_PACKED: *ptr = B->packed.data(); return B->n_reads * B->P.W;
        case ORC_F_NONN_IDS: return put(B->nonN);
        case ORC_F_N_IDS: return put(B->Nids);
        case ORC_F_SORTED_IDS: return put(B->sorted);
        case ORC_F_NEXT_DEDUP: return put(B->next_dedup);
        case ORC_F_PREV_DEDUP: return put(B->prev_dedup);
        case ORC_F_NEXT_OVERLAP: return put(B->next_ov);
        case ORC_F_PREV_OVERLAP: return put(B->prev_ov);
        case ORC_F_OFFSET_OVERLAP: return put(B->off_ov);
        case ORC_F_NEXT_PRUNE: return put(B->next_pr);
        case ORC_F_PREV_PRUNE: return put(B->prev_pr);
        case ORC_F_NEXT: return put(B->next);
        case ORC_F_PREV: return put(B->prev);
        case ORC_F_OFFSET: return put(B->offset);
        case ORC_F_HQ_IDS: return put(B->hq);
        case ORC_F_UNMATCHED_IDS: return put(B->unmatched);
        case ORC_F_HQ_PG: return put(B->hq_pg.text);
        case ORC_F_HQ_REC: return put(B->hq_pg.rec);
        case ORC_F_HQ_PG2: return put(B->hq_pg2);
        case ORC_F_MAP_REC: return put(B->map_rec);
        case ORC_F_MAP_MIS: return put(B->map_mis);
        case ORC_F_UNMAPPED: return put(B->unmapped);
        case ORC_F_UNMAPPED_N: return put(B->unmapped_N);
        case ORC_F_LQ_PG: return put(B->lq_pg.text);
        case ORC_F_LQ_REC: return put(B->lq_pg.rec);
        case ORC_F_NLQ_PG: return put(B->nlq_pg.text);
        case ORC_F_NLQ_REC: return put(B->nlq_pg.rec);
        case ORC_F_LQ_NEXT: return put(B->lq_next);
        case ORC_F_LQ_PREV: return put(B->lq_prev);
        case ORC_F_LQ_OFFSET: return put(B->lq_off);
        case ORC_F_S_STRAND: return put(B->s_strand);
        case ORC_F_S_READ_OFF: return put(B->s_read_off);
        case ORC_F_S_MIS_COUNT: return put(B->s_mis_count);
        case ORC_F_S_MIS_BASE: return put(B->s_mis_base);
        case ORC_F_S_MIS_TABLE: return put(B->s_mis_table);
        case ORC_F_S_LQ_OFF: return put(B->s_lq_off);
        case ORC_F_S_NLQ_OFF: return put(B->s_nlq_off);
        case ORC_F_S_PE_ORDER: return put(B->s_pe_order);
        case ORC_F_S_ID_POS: return put(B->s_id_pos);
        case ORC_F_S_SORTED_REC: return put(B->s_sorted);
        default:
            if (field > ORC_F_S_MIS_OFF && field - ORC_F_S_MIS_OFF < (int)B->s_mis_off.size())
                return put(B->s_mis_off[field - ORC_F_S_MIS_OFF]);
            *ptr = nullptr;
            return 0;
    }
}

// compress.cu:1730-2044: record sort and the stream builders that define the
// block's parity surface (positions, mismatch lists, read order).
int orc_block_streams(orc_block *B, int preserve, int paired) {
    const uint32_t L = B->P.L, W = B->P.W, max_mm = B->P.max_mm;
    const std::string &ref = B->hq_pg.text;
    std::vector<orc_rec> rec = B->hq_pg.rec;
    rec.insert(rec.end(), B->map_rec.begin(), B->map_rec.end());
    std::vector<uint8_t> mis(B->hq_pg.rec.size(), 0);
    mis.insert(mis.end(), B->map_mis.begin(), B->map_mis.end());
    auto gid = [&](const orc_rec &r) -> uint64_t { return (r.word & 1) ? B->Nids[r.read_id] : r.read_id; };
    // compress.cu:1743-1756, stable (R17)
    std::vector<size_t> ord(rec.size());
    std::iota(ord.begin(), ord.end(), (size_t)0);
    if (!preserve)
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return rec_pos(rec[a]) < rec_pos(rec[b]); });
    else
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return gid(rec[a]) < gid(rec[b]); });
    std::vector<orc_rec> R(rec.size());
    std::vector<uint8_t> M(rec.size());
    for (size_t i = 0; i < ord.size(); ++i) {
        R[i] = rec[ord[i]];
        M[i] = mis[ord[i]];
    }
    B->s_sorted = R;
    const size_t n = R.size();
    if (preserve) B->s_id_pos.assign(B->n_reads, 0);
    B->s_strand.assign(n, '0');
    if (!preserve) B->s_read_off.assign(n, 0);
    B->s_pe_order.clear();
    if (!preserve && paired) B->s_pe_order.assign(n, 0);
    B->s_mis_count = M;
    auto width = [](uint32_t m) -> uint64_t { return m == 1 ? 2 : m == 2 ? 3 : (uint64_t)m + 2; };
    std::vector<uint64_t> count_m(max_mm + 1, 0);
    std::vector<uint64_t> off_idx(n, 0);
    uint64_t total_mis = 0;
    for (size_t i = 0; i < n; ++i) {
        if (M[i]) {
            off_idx[i] = width(M[i]) * count_m[M[i]];
            count_m[M[i]]++;
        }
        total_mis += M[i];
    }
    B->s_mis_off.assign(max_mm + 1, {});
    for (uint32_t m = 1; m <= max_mm; ++m) B->s_mis_off[m].assign(width(m) * count_m[m], 0);
    std::string ctx_s(total_mis, 0), base_s(total_mis, 0);
    uint64_t bidx = 0;
    for (size_t i = 0; i < n; ++i) {
        const uint32_t rid = R[i].read_id;
        const uint64_t isN = R[i].word & 1, strand = (R[i].word >> 1) & 1, pos = rec_pos(R[i]);
        const uint32_t m = M[i];
        if (preserve) {
            B->s_id_pos[gid(R[i])] = pos;
        } else {
            if (paired) B->s_pe_order[i] = (uint32_t)gid(R[i]);
            B->s_read_off[i] = (uint16_t)((i == 0) ? 0 : (pos - rec_pos(R[i - 1])));
        }
        B->s_strand[i] = (char)('0' + strand);
        if (!m) continue;
        std::vector<uint16_t> offs;
        for (uint32_t p = 0; p < L; ++p) {
            uint32_t bo = strand == 0 ? p : L - 1 - p;
            char rb = isN ? B->Nascii[(uint64_t)rid * L + bo] : base_at(B->packed.data(), (uint64_t)rid * W, bo);
            if (strand == 1) rb = rev_base(rb);
            if (ref[pos + p] != rb) {
                ctx_s[bidx] = ref[pos + p];
                base_s[bidx] = rb;
                bidx++;
                offs.push_back((uint16_t)p);
            }
        }
        if (offs.size() != m) return -1;  // recount disagrees with the mapping count
        std::vector<uint16_t> &S = B->s_mis_off[m];
        uint64_t k = off_idx[i];
        if (offs[0] < (L - 1 - offs[m - 1])) {  // left to right (compress.cu:1848-1867)
            S[k++] = 0;
            S[k++] = offs[0];
            if (m == 2) {
                S[k++] = (uint16_t)(offs[1] - offs[0] - 1);
            } else if (m > 2) {
                uint32_t mn = 0xFFFFFFFFu;
                for (uint32_t q = 1; q < m; ++q) mn = std::min<uint32_t>(mn, offs[q] - offs[q - 1]);
                S[k++] = (uint16_t)(mn - 1);
                for (uint32_t q = 1; q < m; ++q) S[k++] = (uint16_t)(offs[q] - offs[q - 1] - mn);
            }
        } else {  // right to left (compress.cu:1868-1886)
            S[k++] = 1;
            S[k++] = (uint16_t)(L - 1 - offs[m - 1]);
            if (m == 2) {
                S[k++] = (uint16_t)(offs[1] - offs[0] - 1);
            } else if (m > 2) {
                uint32_t mn = 0xFFFFFFFFu;
                for (uint32_t q = m - 1; q > 0; --q) mn = std::min<uint32_t>(mn, offs[q] - offs[q - 1]);
                S[k++] = (uint16_t)(mn - 1);
                for (uint32_t q = m - 1; q > 0; --q) S[k++] = (uint16_t)(offs[q] - offs[q - 1] - mn);
            }
        }
    }
    // compress.cu:1894-1929: per-context symbol ranks by frequency; ties keep
    // the destination order (std::sort of four elements is an insertion sort).
    static uint64_t cnt[256][256];
    std::memset(cnt, 0, sizeof(cnt));
    for (uint64_t i = 0; i < total_mis; ++i) cnt[(uint8_t)ctx_s[i]][(uint8_t)base_s[i]]++;
    static const char CTX[4] = {'A', 'T', 'C', 'G'};
    static const char DEST[4][4] = {{'C', 'G', 'N', 'T'}, {'A', 'C', 'G', 'N'}, {'A', 'G', 'N', 'T'}, {'A', 'C', 'N', 'T'}};
    uint8_t rank[256][256];
    std::memset(rank, 0, sizeof(rank));
    B->s_mis_table.assign(16, 0);
    for (int c = 0; c < 4; ++c) {
        int idx[4] = {0, 1, 2, 3};
        std::stable_sort(idx, idx + 4, [&](int a, int b) {
            return cnt[(uint8_t)CTX[c]][(uint8_t)DEST[c][a]] > cnt[(uint8_t)CTX[c]][(uint8_t)DEST[c][b]];
        });
        for (int i = 0; i < 4; ++i) rank[(uint8_t)CTX[c]][(uint8_t)DEST[c][idx[i]]] = (uint8_t)i;
        for (int i = 0; i < 4; ++i) B->s_mis_table[c * 4 + rank[(uint8_t)CTX[c]][(uint8_t)DEST[c][i]]] = (uint8_t)DEST[c][i];
    }
    B->s_mis_base.assign(total_mis, 0);
    for (uint64_t i = 0; i < total_mis; ++i) B->s_mis_base[i] = rank[(uint8_t)ctx_s[i]][(uint8_t)base_s[i]];
    // compress.cu:1968-2001: LQ and N-LQ offsets, pair order, id -> position
    const uint64_t hq_len = B->hq_pg.text.size(), lq_len = B->lq_pg.text.size();
    B->s_lq_off.clear();
    B->s_nlq_off.clear();
    uint64_t last = 0;
    for (const auto &r : B->lq_pg.rec) {
        if (!preserve) {
            if (paired) B->s_pe_order.push_back(r.read_id);
            B->s_lq_off.push_back((uint16_t)(L - (rec_pos(r) - last)));
            last = rec_pos(r);
        } else {
            B->s_id_pos[r.read_id] = hq_len + rec_pos(r);
        }
    }
    last = 0;
    for (const auto &r : B->nlq_pg.rec) {
        if (!preserve) {
            if (paired) B->s_pe_order.push_back(B->Nids[r.read_id]);
            B->s_nlq_off.push_back((uint16_t)(L - (rec_pos(r) - last)));
            last = rec_pos(r);
        } else {
            B->s_id_pos[B->Nids[r.read_id]] = hq_len + lq_len + rec_pos(r);
        }
    }
    return 0;
}

void orc_block_free(orc_block *B) { delete B; }

}  // extern "C"
