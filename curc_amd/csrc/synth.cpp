// Seeded synthetic genomes and reads for the CURC workloads.
//
// Every random draw comes from a xoshiro256** stream seeded per genome chunk
// or per read (splitmix64 of seed and index), so the output is the same for
// any thread count.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <cmath>
#include <vector>

#include "../../include/curc_gpu.h"

namespace {

inline uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto& v : s) v = splitmix(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uni() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return n ? (uint64_t)(uni() * (double)n) % n : 0; }
    double normal() {
        double u1 = std::max(uni(), 1e-300), u2 = uni();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

inline uint64_t stream_seed(uint64_t seed, uint64_t tag, uint64_t idx) {
    uint64_t x = seed ^ (tag * 0xD6E8FEB86659FD93ULL);
    splitmix(x);
    x ^= idx * 0x9E3779B97F4A7C15ULL;
    return splitmix(x);
}

const char BASES[4] = {'A', 'C', 'G', 'T'};

inline char draw_base(Rng& r, double gc) {
    double u = r.uni();
    double at = (1.0 - gc) / 2.0, g = gc / 2.0;
    if (u < at) return 'A';
    if (u < at + g) return 'C';
    if (u < at + 2 * g) return 'G';
    return 'T';
}

inline char comp(char c) {
    switch (c) {
        case 'A': return 'T';
        case 'C': return 'G';
        case 'G': return 'C';
        case 'T': return 'A';
        default: return 'N';
    }
}

inline char other_base(Rng& r, char c) {
    int k = 0;
    for (int i = 0; i < 4; ++i)
        if (BASES[i] == c) k = i;
    int o = (int)r.below(3);
    return BASES[(k + 1 + o) & 3];
}

constexpr uint64_t CHUNK = 1 << 16;

}  // namespace

extern "C" {

void curc_synth_default(curc_synth_cfg* c) {
    memset(c, 0, sizeof(*c));
    c->seed = 1;
    c->genome_len = 1000000;
    c->gc = 0.5;
    c->rep_min = 300;
    c->rep_max = 6000;
    c->div_min = 0.05;
    c->div_max = 0.15;
    c->rep_families = 64;
    c->n_reads = 133333;
    c->L = 150;
    c->frag_mean = 400;
    c->frag_sd = 50;
    c->rc_prob = 0.5;
    c->sub_first = 0.005;
    c->sub_last = 0.005;
    c->n_rate = 0.0001;
}

int curc_synth_genome(const curc_synth_cfg* c, char* g) {
    if (!c || !g || c->genome_len == 0) return CURC_EINVAL;
    const uint64_t G = c->genome_len;
    const int64_t nchunks = (int64_t)((G + CHUNK - 1) / CHUNK);
#pragma omp parallel for schedule(static)
    for (int64_t ch = 0; ch < nchunks; ++ch) {
        Rng r(stream_seed(c->seed, 1, (uint64_t)ch));
        uint64_t lo = (uint64_t)ch * CHUNK, hi = std::min(G, lo + CHUNK);
        for (uint64_t i = lo; i < hi; ++i) g[i] = draw_base(r, c->gc);
    }
    if (c->repeat_frac > 0 && c->rep_families > 0) {
        // Repeat families: consensus sequences; copies overwrite the random
        // background at random places until repeat_frac of the genome is covered.
        Rng r(stream_seed(c->seed, 2, 0));
        std::vector<std::string> fam(c->rep_families);
        for (auto& f : fam) {
            uint64_t len = c->rep_min + r.below((uint64_t)c->rep_max - c->rep_min + 1);
            f.resize(len);
            for (auto& b : f) b = draw_base(r, c->gc);
        }
        uint64_t target = (uint64_t)(c->repeat_frac * (double)G);
        uint64_t covered = 0;
        uint64_t guard = 0;
        while (covered < target && guard++ < 100000000ULL) {
            const std::string& f = fam[r.below(fam.size())];
            if (f.size() >= G) break;
            uint64_t pos = r.below(G - f.size());
            double div = c->div_min + (c->div_max - c->div_min) * r.uni();
            bool rc = r.uni() < 0.5;
            for (uint64_t j = 0; j < f.size(); ++j) {
                char b = rc ? comp(f[f.size() - 1 - j]) : f[j];
                if (r.uni() < div) b = other_base(r, b);
                g[pos + j] = b;
            }
            covered += f.size();
        }
    }
    return CURC_OK;
}

int curc_synth_reads(const curc_synth_cfg* c, const char* g, char* out) {
    if (!c || !g || !out || c->L == 0 || c->genome_len < c->L) return CURC_EINVAL;
    const uint64_t G = c->genome_len;
    const uint32_t L = c->L;
    if (c->paired && (c->n_reads % 2)) return CURC_EINVAL;
    const int64_t units = (int64_t)(c->paired ? c->n_reads / 2 : c->n_reads);
#pragma omp parallel for schedule(static, 4096)
    for (int64_t u = 0; u < units; ++u) {
        Rng r(stream_seed(c->seed, 3, (uint64_t)u));
        char* dst = out + (uint64_t)u * L * (c->paired ? 2 : 1);
        auto finish = [&](char* rd) {
            for (uint32_t i = 0; i < L; ++i) {
                double p = c->sub_first;
                if (L > 1) p += (c->sub_last - c->sub_first) * (double)i / (double)(L - 1);
                if (r.uni() < p) rd[i] = other_base(r, rd[i]);
                if (c->n_rate > 0 && r.uni() < c->n_rate) rd[i] = 'N';
            }
        };
        if (!c->paired) {
            uint64_t s = r.below(G - L + 1);
            bool rc = r.uni() < c->rc_prob;
            for (uint32_t i = 0; i < L; ++i) dst[i] = rc ? comp(g[s + L - 1 - i]) : g[s + i];
            finish(dst);
        } else {
            double fl = c->frag_mean + c->frag_sd * r.normal();
            uint64_t f = (uint64_t)std::max((double)L, std::min(fl, (double)G));
            uint64_t s = r.below(G - f + 1);
            bool flip = r.uni() < 0.5;
            char* r1 = dst;
            char* r2 = dst + L;
            // r1 reads the fragment's 5' end forward, r2 its 3' end reverse
            // complemented; a flipped fragment comes from the other strand.
            for (uint32_t i = 0; i < L; ++i) {
                if (!flip) {
                    r1[i] = g[s + i];
                    r2[i] = comp(g[s + f - 1 - i]);
                } else {
                    r1[i] = comp(g[s + f - 1 - i]);
                    r2[i] = g[s + i];
                }
            }
            finish(r1);
            finish(r2);
        }
    }
    return CURC_OK;
}

// FASTQ records of n reads (read i at reads + i * stride) with the layout of
// curc_amd.fastq.fastq_bytes: "@<prefix><id, 10 digits>[/mate]\n", the
// sequence, "\n+\n", L quality bytes 'I', "\n".  Records have a fixed
// length, so each is written at its own offset (in parallel).
uint64_t curc_synth_fastq_record_len(uint32_t L, const char* prefix, int mate) {
    return 1 + strlen(prefix) + 10 + (mate > 0 ? 2 : 0) + 1 + L + 3 + L + 1;
}

int curc_synth_fastq(const char* reads, uint64_t n, uint32_t L, uint64_t stride, uint64_t first_id, const char* prefix,
                     int mate, char* out) {
    if ((!reads && n) || (!out && n) || !prefix || L == 0 || stride < L || mate < 0 || mate > 9) return CURC_EINVAL;
    const uint64_t plen = strlen(prefix);
    const uint64_t rec = curc_synth_fastq_record_len(L, prefix, mate);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        char* o = out + (uint64_t)i * rec;
        *o++ = '@';
        memcpy(o, prefix, plen);
        o += plen;
        uint64_t id = first_id + (uint64_t)i;
        for (int d = 9; d >= 0; --d) {
            o[d] = (char)('0' + id % 10);
            id /= 10;
        }
        o += 10;
        if (mate > 0) {
            *o++ = '/';
            *o++ = (char)('0' + mate);
        }
        *o++ = '\n';
        memcpy(o, reads + (uint64_t)i * stride, L);
        o += L;
        *o++ = '\n';
        *o++ = '+';
        *o++ = '\n';
        memset(o, 'I', L);
        o += L;
        *o = '\n';
    }
    return CURC_OK;
}

}  // extern "C"
