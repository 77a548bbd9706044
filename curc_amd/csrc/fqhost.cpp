// FASTQ record parse on the host (see fqhost.h).
#include "fqhost.h"

#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>

namespace curc {
namespace {

// Bit j set for each byte B[base + j] == '\n', j < 64, base + j < nb.
inline uint64_t nl_mask_scalar(const uint8_t* B, uint64_t base, uint64_t nb) {
    uint64_t m = 0;
    const uint64_t e = std::min<uint64_t>(nb - base, 64);
    for (uint64_t j = 0; j < e; ++j)
        if (B[base + j] == '\n') m |= 1ull << j;
    return m;
}

__attribute__((target("avx2"))) inline uint64_t nl_mask_avx2(const uint8_t* p) {
    const __m256i nl = _mm256_set1_epi8('\n');
    const uint32_t lo = (uint32_t)_mm256_movemask_epi8(_mm256_cmpeq_epi8(_mm256_loadu_si256((const __m256i*)p), nl));
    const uint32_t hi =
        (uint32_t)_mm256_movemask_epi8(_mm256_cmpeq_epi8(_mm256_loadu_si256((const __m256i*)(p + 32)), nl));
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

// Whole 64-byte blocks without vector instructions (8 bytes at a time,
// exact zero-byte test).
inline uint64_t nl_mask_swar(const uint8_t* p) {
    uint64_t m = 0;
    for (int k = 0; k < 8; ++k) {
        uint64_t v;
        memcpy(&v, p + 8 * k, 8);
        const uint64_t x = v ^ 0x0A0A0A0A0A0A0A0Aull;
        const uint64_t z = ~(((x & 0x7F7F7F7F7F7F7F7Full) + 0x7F7F7F7F7F7F7F7Full) | x | 0x7F7F7F7F7F7F7F7Full);
        // z has bit 7 of each byte set where the byte was '\n'
        uint64_t b = 0;
        for (int j = 0; j < 8; ++j) b |= ((z >> (8 * j + 7)) & 1ull) << j;
        m |= b << (8 * k);
    }
    return m;
}

// One thread's part: the record chain from record start p (line phase 0)
// up to the first record start at or past `stop`, or to the end of the
// bytes.  Sequences of valid records go to o (at most cap of them).
struct Walk {
    uint64_t end = 0;     // first record start >= stop, or the start after the last complete record
    uint64_t recs = 0;    // complete records
    uint64_t copied = 0;  // sequences written
    uint32_t phase = 0;   // newlines since the last record start (mod 4) where the bytes ended
    bool len_err = false;
    bool overflow = false;
};

template <bool AVX2>
void walk(const uint8_t* B, uint64_t nb, uint64_t p, uint64_t stop, uint32_t L, uint8_t* o, uint64_t cap, Walk* w) {
    Walk r;
    r.end = p;
    if (p >= stop || p >= nb) {
        *w = r;
        return;
    }
    uint32_t phase = 0;
    uint64_t seq = 0;
    uint64_t base = p & ~63ull;
    uint64_t m0 = ~0ull << (p - base);
    for (; base < nb; base += 64) {
        uint64_t m;
        if (base + 64 <= nb)
            m = AVX2 ? nl_mask_avx2(B + base) : nl_mask_swar(B + base);
        else
            m = nl_mask_scalar(B, base, nb);
        m &= m0;
        m0 = ~0ull;
        while (m) {
            const uint64_t pos = base + (uint64_t)__builtin_ctzll(m);
            m &= m - 1;
            if (phase == 0) {
                seq = pos + 1;
                phase = 1;
            } else if (phase == 1) {
                if (pos - seq != L) {
                    r.len_err = true;
                } else if (r.copied < cap) {
                    memcpy(o + r.copied * L, B + seq, L);
                    ++r.copied;
                } else {
                    r.overflow = true;
                }
                phase = 2;
            } else if (phase == 2) {
                phase = 3;
            } else {
                phase = 0;
                ++r.recs;
                r.end = pos + 1;
                if (r.end >= stop) {
                    *w = r;
                    return;
                }
            }
        }
    }
    r.phase = phase;
    *w = r;
}

bool has_avx2() {
    static const bool v = __builtin_cpu_supports("avx2");
    return v;
}

void walk_any(const uint8_t* B, uint64_t nb, uint64_t p, uint64_t stop, uint32_t L, uint8_t* o, uint64_t cap,
              Walk* w) {
    if (has_avx2())
        walk<true>(B, nb, p, stop, L, o, cap, w);
    else
        walk<false>(B, nb, p, stop, L, o, cap, w);
}

// First byte at or after q that is '\n' (nb if none).
uint64_t next_nl(const uint8_t* B, uint64_t nb, uint64_t q) {
    if (q >= nb) return nb;
    const void* f = memchr(B + q, '\n', nb - q);
    return f ? (uint64_t)((const uint8_t*)f - B) : nb;
}

// A thread's start guess: the first line start p >= b with '@' at p, a next
// line of exactly L bytes and a line starting with '+' after it; nb if none.
uint64_t guess_start(const uint8_t* B, uint64_t nb, uint64_t b, uint32_t L) {
    uint64_t q = b - 1;  // b >= 1: find the newline that ends the line before p
    for (;;) {
        const uint64_t nl = next_nl(B, nb, q);
        const uint64_t p = nl + 1;
        if (p >= nb) return nb;
        q = p;
        if (B[p] != '@') continue;
        const uint64_t h = next_nl(B, nb, p);
        if (h >= nb) return nb;
        const uint64_t s = next_nl(B, nb, h + 1);
        if (s >= nb) return nb;
        if (s - (h + 1) == L && s + 1 < nb && B[s + 1] == '+') return p;
    }
}

uint64_t seg_cap(uint64_t bytes, uint32_t L) { return bytes / ((uint64_t)L + 4) + 1; }

}  // namespace

uint64_t fq_host_cap(uint64_t nb, uint32_t L, uint32_t T) {
    return (nb / ((uint64_t)L + 4) + (uint64_t)FQH_RANGES * std::max<uint32_t>(T, 1) + 1) * (uint64_t)L;
}

void fq_host_parse(const uint8_t* B, uint64_t nb, uint32_t L, uint32_t T, uint8_t* out, FqHostParse* res,
                   const FqEmit& emit) {
    FqHostParse r;
    if (nb == 0) {
        *res = r;
        return;
    }
    if (B[0] != '@') {
        r.err = FQH_START;
        *res = r;
        return;
    }
    T = std::max<uint32_t>(T, 1);
    const uint32_t R = FQH_RANGES * T;
    std::vector<uint64_t> S(R + 1);
    S[0] = 0;
    S[R] = nb;
    for (uint32_t q = 1; q < R; ++q) S[q] = std::max(S[q - 1], guess_start(B, nb, nb * q / R, L));
    std::vector<uint64_t> off(R + 1, 0), cap(R);
    for (uint32_t q = 0; q < R; ++q) {
        cap[q] = seg_cap(S[q + 1] - S[q], L);
        off[q + 1] = off[q] + cap[q] * L;
    }
    std::vector<Walk> w(R);
    // T workers take the ranges in order; this thread hands each range's
    // sequences on as soon as it and every range before it are done
    std::mutex mu;
    std::condition_variable cv;
    std::vector<char> done(R, 0);
    std::atomic<uint32_t> next{0};
    auto worker = [&]() {
        for (;;) {
            const uint32_t q = next.fetch_add(1);
            if (q >= R) return;
            walk_any(B, nb, S[q], S[q + 1], L, out + off[q], cap[q], &w[q]);
            {
                std::lock_guard<std::mutex> lk(mu);
                done[q] = 1;
            }
            cv.notify_all();
        }
    };
    std::vector<std::thread> th;
    th.reserve(T);
    for (uint32_t t = 0; t < T; ++t) th.emplace_back(worker);
    std::exception_ptr failed;
    uint64_t base = 0;
    bool len_err = false;
    auto hand_on = [&](uint64_t o, uint64_t count) {
        if (!count) return;
        const FqSeg g{o, count};
        r.segs.push_back(g);
        if (emit && !failed) {
            try {
                emit(g, base);
            } catch (...) {
                failed = std::current_exception();
            }
        }
        base += count;
    };
    int broken = -1;  // first range whose successor's guess the chain did not land on
    bool overflow = false;
    uint32_t last = 0;
    for (uint32_t q = 0; q < R; ++q) {
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return done[q] != 0; });
        }
        if (w[q].overflow) {
            broken = (int)q;
            overflow = true;
            break;
        }
        hand_on(off[q], w[q].copied);
        r.n += w[q].recs;
        len_err |= w[q].len_err;
        if (S[q] < nb) last = q;
        if (q + 1 < R && S[q + 1] < nb && w[q].end != S[q + 1]) {
            broken = (int)q;
            break;
        }
    }
    next.store(R);  // no more ranges for the workers
    for (auto& x : th) x.join();
    Walk tail;
    const Walk* lw = &w[last];
    if (broken >= 0) {
        // a start guess was not a record start (or, never expected, a range
        // overflowed): the rest of the bytes in one pass from the last record
        // start the chain confirmed, into the space of the ranges after it
        r.resync = true;
        const uint32_t q = (uint32_t)broken;
        const uint32_t q1 = overflow ? q : q + 1;
        uint64_t tcap = 0;
        for (uint32_t x = q1; x < R; ++x) tcap += cap[x];
        walk_any(B, nb, overflow ? S[q] : w[q].end, nb, L, out + off[q1], tcap, &tail);
        hand_on(off[q1], tail.copied);
        r.n += tail.recs;
        len_err |= tail.len_err;
        lw = &tail;
    }
    if (failed) std::rethrow_exception(failed);
    if (B[nb - 1] != '\n' || lw->phase != 0 || lw->end != nb) {
        r.err = FQH_TRUNC;
    } else if (len_err) {
        r.err = FQH_LEN;
    }
    if (r.err) r.segs.clear();
    *res = r;
}

}  // namespace curc

// ---- C ABI: the host parse on its own (include/curc_gpu.h) ----
#include "../../include/curc_gpu.h"

extern "C" {

uint64_t curc_fastq_host_cap(uint64_t bytes, uint32_t L, uint32_t threads) {
    return curc::fq_host_cap(bytes, L, threads);
}

int curc_fastq_host_parse(const char* fq, uint64_t bytes, uint32_t L, uint32_t threads, char* out,
                          uint64_t out_bytes, uint64_t* n_reads, int* format_error) {
    if ((!fq && bytes) || !n_reads || !format_error || L == 0 || threads == 0 ||
        out_bytes < curc::fq_host_cap(bytes, L, threads) || (!out && out_bytes))
        return CURC_EINVAL;
    curc::FqHostParse r;
    curc::fq_host_parse(reinterpret_cast<const uint8_t*>(fq), bytes, L, threads, reinterpret_cast<uint8_t*>(out), &r);
    *format_error = r.err;
    *n_reads = r.err ? 0 : r.n;
    if (r.err) return CURC_EINVAL;
    // segments move down into one run, in record order
    uint64_t at = 0;
    for (const auto& s : r.segs) {
        if (s.off != at) memmove(out + at, out + s.off, s.count * L);
        at += s.count * L;
    }
    return CURC_OK;
}

}  // extern "C"
