/*
 * RefMatcher restatement (compress.cu:597-1001) -- TEST INFRASTRUCTURE ONLY
 * (see oracle.h).
 *
 * The source text is the HQ pseudogenome; a destination text (the HQ
 * pseudogenome itself, the LQ or the N-read LQ pseudogenome) is reverse
 * complemented, probed at every third position against a k=36 maRushPrime1
 * index of the source (every fifth position), and each verified seed is
 * extended both ways.  Matches of at least 50 bases are replaced by '%' in
 * the destination and described by the MapOff / MapLen streams.
 *
 * Deterministic choices where the reference is racy:
 *   - buckets hold the first 12 source positions by ascending index, visited
 *     in descending index (the sequential meaning of compress.cu:713-741);
 *   - results are collected in segment (binary destination) or position
 *     (character destination) order and truncated to the capacity there
 *     (compress.cu:794-797, 843-846);
 *   - probe positions whose 36-mer runs past the destination's end (possible
 *     only in the last segment of a binary destination) are skipped: the
 *     reference reads past the text there and any match it records extends
 *     beyond the destination.
 * Texts shorter than 36 bases yield no matches (the reference's unsigned
 * position count wraps there).
 */
#include <algorithm>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "oracle.h"

namespace {

constexpr uint64_t RM_K = 36;              // compress.cu:678 kmer_size
constexpr uint64_t RM_SRC_STEP = 5;        // compress.cu:676 src_step
constexpr uint64_t RM_DEST_STEP = 3;       // compress.cu:677 dest_step
constexpr uint64_t RM_TARGET = 50;         // compress.cu:679 target_match_len
constexpr uint32_t RM_BUCKET = 12;         // compress.cu:680 bucket_size_limit
constexpr double RM_REF_RATIO = 0.2;       // compress.cu:681
constexpr double RM_UNMAP_RATIO = 0.21;    // compress.cu:682
constexpr uint64_t RM_SEG = 4;             // compress.cu:750 dest_segement_step

// third_party/lzma/VarLenDNACoder/helper.cpp:253-268 (complementsLut)
struct CompLut {
    char t[256];
    CompLut() {
        std::memset(t, 0, sizeof(t));
        const char *from = "AaCcGgTtNnUuYyRrKkMmBbDdHhVv";
        const char *to = "TTGGCCAANNAARRYYMMKKVVHHDDBB";
        for (int i = 0; from[i]; ++i) t[(uint8_t)from[i]] = to[i];
    }
};
const CompLut COMP;

// compress.cu:617-629 (maRushPrime1HashChar<36>): 4 characters at a time,
// read as a little-endian u32.
uint32_t marush36(const char *s) {
    uint64_t h = RM_K;
    for (uint32_t j = 0; j < RM_K / 4; ++j) {
        uint32_t k = (uint32_t)(uint8_t)s[4 * j] | (uint32_t)(uint8_t)s[4 * j + 1] << 8 |
                     (uint32_t)(uint8_t)s[4 * j + 2] << 16 | (uint32_t)(uint8_t)s[4 * j + 3] << 24;
        k += j;
        h ^= k;
        h *= 171717ull;
    }
    return (uint32_t)h;
}

// helper.cpp:179-187
void write_frugal(std::string &o, uint64_t v) {
    while (v >= 128) {
        o.push_back((char)(128 + (v % 128)));
        v /= 128;
    }
    o.push_back((char)v);
}

struct Match {
    uint64_t src, len, dest;
    bool operator<(const Match &o) const {  // compress.cu:650-660
        if (dest != o.dest) return dest < o.dest;
        if (src != o.src) return src < o.src;
        return len < o.len;
    }
    bool operator==(const Match &o) const { return src == o.src && len == o.len && dest == o.dest; }
};

}  // namespace

struct orc_refmatch {
    std::string text, map_off, map_len;
    std::vector<uint64_t> found;    // (src, len, dest) per result, collection order, before capacity
    std::vector<uint64_t> matches;  // (src, len, dest) after transform, sort and unique
    uint64_t capacity = 0;
};

extern "C" {

uint32_t orc_marush36(const char *s) { return marush36(s); }

orc_refmatch *orc_ref_match(const char *src, uint64_t src_len, const char *dest, uint64_t dest_len,
                            int dest_contain_N, int dest_is_ref) {
    orc_refmatch *R = new orc_refmatch();
    // compress.cu:873, 953-954: the destination is reverse complemented in
    // place and back again, so its text leaves as comp(comp(c)) per symbol.
    std::string rc(dest_len, '\0');
    for (uint64_t i = 0; i < dest_len; ++i) rc[i] = COMP.t[(uint8_t)dest[dest_len - 1 - i]];
    R->text.resize(dest_len);
    for (uint64_t i = 0; i < dest_len; ++i) R->text[i] = COMP.t[(uint8_t)rc[dest_len - 1 - i]];
    if (dest_len < RM_K || src_len < RM_K) return R;

    // RefMatcher::init (compress.cu:698-741)
    const uint64_t index_count = (src_len - RM_K + 1 + RM_SRC_STEP - 1) / RM_SRC_STEP;
    const uint32_t H = orc_hash_size(index_count);
    const uint32_t mask = H - 1;
    std::vector<uint32_t> hs(index_count);
    std::vector<uint8_t> cnt(H, 0);
    for (uint64_t i = 0; i < index_count; ++i) {
        hs[i] = marush36(src + i * RM_SRC_STEP) & mask;
        if (cnt[hs[i]] < RM_BUCKET) ++cnt[hs[i]];
    }
    std::vector<uint32_t> range((uint64_t)H + 1, 0);
    for (uint64_t h = 0; h < H; ++h) range[h + 1] = range[h] + cnt[h];
    // bucket h holds its first cnt[h] occurrences at range[h] + 0.. in
    // ascending index order; probes visit them from the last
    std::vector<uint32_t> bucket(range[H]);
    std::vector<uint8_t> fillc(H, 0);
    for (uint64_t i = 0; i < index_count; ++i) {
        const uint32_t h = hs[i];
        if (fillc[h] < cnt[h]) bucket[range[h] + fillc[h]++] = (uint32_t)i;
    }

    // one probe (compress.cu:759-803 binary, :815-851 character)
    const bool binary = !dest_contain_N;
    auto probe = [&](uint64_t dp, Match &m) -> bool {
        const uint32_t h = marush36(&rc[dp]) & mask;
        for (int q = (int)cnt[h] - 1; q >= 0; --q) {
            const uint64_t sp = (uint64_t)bucket[range[h] + q] * RM_SRC_STEP;
            if (binary && dest_is_ref && (dest_len - sp < dp)) continue;
            uint64_t p1 = dp + RM_K - 1, p2 = sp + RM_K - 1;
            while (++p1 != dest_len && ++p2 != src_len && rc[p1] == src[p2]) {
            }
            const uint64_t right = p1;
            p1 = dp;
            p2 = sp;
            while (p1 != 0 && p2 != 0 && rc[--p1] == src[--p2]) {
            }
            if (right - p1 > RM_TARGET) {
                if (std::memcmp(&rc[dp], src + sp, RM_K) == 0) {
                    m = Match{p2 + 1, right - p1 - 1, p1 + 1};
                    return true;
                }
            }
        }
        return false;
    };

    const uint64_t dic = (dest_len - RM_K + 1 + RM_DEST_STEP - 1) / RM_DEST_STEP;
    R->capacity = (uint64_t)((double)dic * (dest_is_ref ? RM_REF_RATIO : RM_UNMAP_RATIO));
    std::vector<Match> found;
    Match m{};
    if (binary) {
        const uint64_t segs = (dic + RM_SEG - 1) / RM_SEG;
        for (uint64_t g = 0; g < segs; ++g)
            for (uint64_t s = 0; s < RM_SEG; ++s) {
                const uint64_t dp = (g * RM_SEG + s) * RM_DEST_STEP;
                if (dp + RM_K > dest_len) continue;
                if (probe(dp, m)) {
                    found.push_back(m);
                    break;  // compress.cu:798 (goto finish)
                }
            }
    } else {
        for (uint64_t i = 0; i < dic; ++i)
            if (probe(i * RM_DEST_STEP, m)) found.push_back(m);
    }
    for (const Match &f : found) {
        R->found.push_back(f.src);
        R->found.push_back(f.len);
        R->found.push_back(f.dest);
    }
    if (found.empty()) return R;  // compress.cu:908-913: empty streams
    if (found.size() > R->capacity) found.resize(R->capacity);

    // compress.cu:914-932
    for (Match &f : found) {
        f.dest = dest_len - (f.dest + f.len);
        if (dest_is_ref) {
            if (f.src > f.dest) std::swap(f.src, f.dest);
            const uint64_t end_src = f.src + f.len;
            if (end_src > f.dest) {
                const uint64_t margin = (end_src - f.dest + 1) / 2;
                f.len -= margin;
                f.dest += margin;
            }
        }
    }
    std::sort(found.begin(), found.end());
    found.erase(std::unique(found.begin(), found.end()), found.end());
    for (const Match &f : found) {
        R->matches.push_back(f.src);
        R->matches.push_back(f.len);
        R->matches.push_back(f.dest);
    }

    // compress.cu:953-993: '%' rewrite and the MapOff / MapLen streams
    const bool std_len = src_len <= 0xFFFFFFFFull;
    write_frugal(R->map_len, RM_TARGET);
    std::string out;
    out.reserve(dest_len);
    uint64_t pos = 0;
    for (Match f : found) {
        if (f.dest < pos) {
            const uint64_t overflow = pos - f.dest;
            if (overflow >= f.len) continue;
            f.len -= overflow;
            f.dest += overflow;
        }
        if (f.len < RM_TARGET) continue;
        out.append(R->text, pos, f.dest - pos);
        out.push_back('%');
        if (std_len) {
            const uint32_t v = (uint32_t)f.src;
            R->map_off.append((const char *)&v, 4);
        } else {
            R->map_off.append((const char *)&f.src, 8);
        }
        write_frugal(R->map_len, f.len - RM_TARGET);
        pos = f.dest + f.len;
    }
    out.append(R->text, pos, std::string::npos);
    R->text.swap(out);
    return R;
}

uint64_t orc_refmatch_get(orc_refmatch *R, int field, const void **ptr) {
    switch (field) {
        case ORC_RM_TEXT: *ptr = R->text.data(); return R->text.size();
        case ORC_RM_MAP_OFF: *ptr = R->map_off.data(); return R->map_off.size();
        case ORC_RM_MAP_LEN: *ptr = R->map_len.data(); return R->map_len.size();
        case ORC_RM_FOUND: *ptr = R->found.data(); return R->found.size() / 3;
        case ORC_RM_MATCHES: *ptr = R->matches.data(); return R->matches.size() / 3;
        case ORC_RM_CAPACITY: *ptr = &R->capacity; return 1;
        default: *ptr = nullptr; return 0;
    }
}

void orc_refmatch_free(orc_refmatch *R) { delete R; }

}  // extern "C"
