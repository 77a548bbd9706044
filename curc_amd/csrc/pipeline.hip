// Block driver and C ABI of libcurc_gpu (see include/curc_gpu.h).
//
// One block runs as a single stream of kernels; the host is consulted only
// where a size decides an allocation (after the N split, after dedup, after
// the partition, inside assembly, after mapping).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <chrono>
#include <numeric>
#include <cstdlib>
#include <exception>
#include <thread>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/curc_gpu.h"
#include "assembly.cuh"
#include "decode.cuh"
#include "fastq.cuh"
#include "kernels.cuh"
#include "mapphase.cuh"
#include "mlp.cuh"
#include "refmatch.cuh"
#include "graph_cache.h"
#include "host_pool.h"
#include "fqhost.h"
#include "streams.cuh"

using namespace curc;

// Fixed device scratch of curc_fastq_pack_static, for callers that cannot
// allocate: counters in the first 256 bytes, then the parse's tile counts.
// Each GPU holds its own copy for the life of the process; its address on
// a device is looked up once (static_scratch_init) and used from then on.
static constexpr uint64_t STATIC_SCRATCH_WORDS = 1ull << 20;  // 8 MiB
static constexpr uint64_t STATIC_CNT_WORDS = 32;               // 256 bytes of u32 counters
static __device__ __align__(256) uint64_t scratch[STATIC_SCRATCH_WORDS];

static constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static uint64_t* g_scratch_ptr[MAX_DEVICES] = {};

// Resolves `scratch` on the current device (once per device); returns the
// device address, or null when the lookup fails.
static uint64_t* static_scratch_init(int device) {
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[device] = static_cast<uint64_t*>(p);
    }
    return g_scratch_ptr[device];
}

namespace {

// Waits for a stream's work by polling it (the block's many short host
// waits: counts read back between stages).  A blocking wait wakes the host
// thread tens of microseconds after the GPU has finished, and the GPU idles
// meanwhile; polling returns within a few.
#ifndef CURC_SPIN_SYNC
#define CURC_SPIN_SYNC 1
#endif
inline hipError_t stream_wait(hipStream_t st) {
#if CURC_SPIN_SYNC
    hipError_t e;
    while ((e = hipStreamQuery(st)) == hipErrorNotReady) {
    }
    return e;
#else
    return hipStreamSynchronize(st);
#endif
}

enum Slot {
    S_ASCII, S_PACKED, S_ISN, S_NONN, S_NIDS, S_NASCII, S_CNT,
    S_K0, S_K1, S_V0, S_V1, S_NEXT, S_PREV, S_OFF,
    S_VA, S_VK, S_KR, S_SL0, S_SL1, S_CHOICE, S_GRP, S_OWNER, S_CONT, S_HEAD, S_MEMB, S_RESL,
    S_PL0, S_PL1, S_PK0, S_PK1, S_RNG, S_RES, S_HQ, S_UNM, S_PRE, S_SUF,
    S_PG_HQ, S_PG_LQ, S_PGIDS, S_PGPOS, S_PGNB, S_SORTED,
    S_A_RANK, S_A_NXT, S_A_PRV, S_A_OFF, S_A_ST0, S_A_ST1, S_A_MEMB, S_A_CKEY, S_A_CLEN, S_A_CCNT,
    S_A_HK0, S_A_HK1, S_A_HV0, S_A_HV1, S_A_OLEN, S_A_OCNT, S_A_CPOS, S_A_CREC,
    S_MK0, S_MK1, S_MV0, S_MV1, S_KRF, S_MKR, S_MPOS, S_MBK, S_PSLOT, S_PFILT, S_SUFW, S_KPOOL, S_TPOS, S_TSID, S_TP0, S_TP1,
    S_THEAD, S_TRUN, S_TK0, S_TK1, S_LAST, S_WORD,
    S_MREC_ID, S_MREC_W, S_MIS, S_UNMAPPED, S_UNMAPPED_N, S_RECAOS,
    S_LQPACKED, S_LQNEXT, S_LQPREV, S_LQOFF,
    S_N3, S_NNEXT, S_NPREV, S_NOFF, S_NSORT, S_NPG, S_POOL,
    S_LQIDS, S_LQPOS, S_LQNB, S_NLQIDS, S_NLQPOS, S_NLQNB,
    S_ST_RID, S_ST_RWORD, S_ST_STRAND, S_ST_MISC, S_ST_ROFF, S_ST_PE, S_ST_IDPOS, S_ST_MKEY, S_ST_MIDX,
    S_ST_MKEY2, S_ST_MIDX2, S_ST_BIDX, S_ST_OIDX, S_ST_GSTART, S_ST_MBASE, S_ST_PAIR, S_ST_OFFS, S_ST_HIST,
    S_ST_RANK, S_ST_SYM, S_ST_LQOFF, S_ST_NLQOFF, S_MWORK, S_SLINK, S_PLINK,
    S_RM_K0, S_RM_K1, S_RM_V0, S_RM_V1, S_RM_KRF, S_RM_KR, S_RM_POS, S_RM_RC, S_RM_S, S_RM_L, S_RM_D,
    S_RM_FLAG, S_RM_FS, S_RM_FL, S_RM_FD, S_RM_SK0, S_RM_SK1, S_RM_SV0, S_RM_SV1, S_RM_US, S_RM_UL, S_RM_UD,
    S_RM_T, S_RM_TMAX, S_RM_J0, S_RM_J1, S_RM_MARK, S_RM_AS, S_RM_AL, S_RM_AD, S_RM_ASTART, S_RM_AEND,
    S_RM_LB, S_RM_GAP, S_RM_MOFF, S_RM_MLEN, S_RM_TEXT, S_RM_WORK,
    S_MACT0, S_MACT1, S_MSEEN, S_MCNT, S_MSTOP, S_MX, S_MLOGIT, S_TRX, S_TRY, S_TRLAST, S_TRSLOT, S_TRCNT,
    S_OVCAND, S_OVNC, S_LB, S_OVVB, S_RECHQ, S_RECMAP, S_MK2, S_MK3, S_MV2, S_MV3, S_KRF2,
    S_FQ0, S_FQ1, S_FQSEQ, S_FQTILE,
    S_D_TEXT, S_D_STRAND, S_D_CNT, S_D_BASE, S_D_STREAMS, S_D_SSTART, S_D_TAB, S_D_ROFF, S_D_LQ, S_D_NLQ,
    S_D_PE, S_D_IDPOS, S_D_PV, S_D_PX, S_D_HQV, S_D_HQX, S_D_BV, S_D_BX, S_D_K0, S_D_K1, S_D_V0, S_D_V1,
    S_D_GS, S_D_FLAT, S_D_OUT1, S_D_OUT2, S_OVPASS, S_WORK, S_FMNF, S_FMVB, S_FMFILT, S_FMLEAD, S_FMGC, S_FMGM,
    S_NMASK,
    S_OVLIST,
    S_A_RS0, S_A_RS1, S_A_RUL, S_A_RIDX, S_A_RLIST, S_DEQ,
    S_LAST_NAMED,
    S_PRIM = 240  // primitives use slots 240.. (scan recursion, sort hist, compaction)
};

// Device counters (u32) in slot S_CNT.
static_assert(S_LAST_NAMED <= S_PRIM, "scratch slots overlap");

enum Cnt {
    C_N, C_NONN, C_NN, C_NPRE, C_NSUF, C_NS0, C_NS1, C_RESN, C_NP0, C_NP1,
    C_HQ, C_UNM, C_PRE, C_SUF, C_NMAP, C_NUNMAP, C_NUNMAPN, C_TMP, C_ASM, C_POOL, C_RM, C_MA0, C_MA1, C_TIE,
    C_FQERR,
    C_COUNT = 32
};
static_assert(C_COUNT * 4 <= STATIC_CNT_WORDS * 8, "counters exceed the fixed scratch's counter area");
static_assert(CURC_FASTQ_STATIC_MAX_BYTES / FQ_TILE + 2 <= STATIC_SCRATCH_WORDS - STATIC_CNT_WORDS,
              "tile counts of the largest accepted input exceed the fixed scratch");

inline uint32_t grid_for(uint64_t n, uint32_t cap = 8192) {
    uint64_t g = (n + 255) / 256;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (uint32_t)g;
}

inline uint32_t hash_size(uint64_t n) {  // util.cuh:72-80
    uint32_t hs;
    uint32_t i = 24;
    do {
        hs = 1u << (i++);
    } while (i <= 31 && hs < n);
    return hs;
}

inline uint32_t log2u(uint64_t v) {
    uint32_t r = 0;
    while ((1ull << r) < v) ++r;
    return r;
}

struct StageStat {
    std::string name;
    double ms = 0;
    uint64_t launches = 0;
    double alg_bytes = 0;
    double items = 0;
};

struct PendingTimer {
    int stat;
    hipEvent_t a, b;
};

}  // namespace

// Launch shapes that change speed only, never results (set by
// curc_ctx_set_tuning / the autotuner; CURC_TUNE="key=value,..." at context
// creation overrides them).
// Defaults: the autotuner's picks on a 20M x 150 bp block (MI355X).
struct Tuning {
    uint32_t ov_grid = 16384;    // workgroups (cap) of the overlap probe and claim finalisation
    uint32_t fm_grid = 4096;     // workgroups (cap) of the exact-matching probe and claim finalisation
    uint32_t map_grid = 32768;   // workgroups (cap) of the mapping kernels
    uint32_t radix_items = 16;   // keys per thread of a radix-sort tile (8, 16, 32)
    uint32_t filter_bits = 32;   // prefix k-mer filter bits per prefix (8 .. 128, power of two)
    uint32_t fm_fused = 4096;    // exact-matching lists up to this bound run in one fused launch (0: never)
    uint32_t map_both = 2;       // both strand passes in one launch on a two-strand index of 128-byte lines (1) or 32-byte slots (2)
    uint32_t ov_queue = 3;       // overlap probe: 0 one suffix per thread, 1 queued in LDS, 2 queued with verdicts,
                                 // 3 verdict words against a linked-suffix bitmap
    uint32_t workers = 4;        // streams of curc_batch_run when the caller passes 0
    uint32_t fm_static = 1;      // exact matching over lists kept in place (1) or compacted per offset (0)
    uint32_t fq_host = 8;        // host threads that parse FASTQ records before the copy (0: parse on the device)
};

struct curc_ctx {
    int device = 0;
    uint32_t cus = 256;  // compute units of the device (hipDeviceProp_t::multiProcessorCount)
    Tuning tune;
    hipStream_t st = nullptr;
    // copies of results to pinned host memory, ordered after the compute
    // stream's work by the events cev, so that later stages run meanwhile;
    // joined before a block is returned (curc_block_run).  Created on first
    // use, and not used by the contexts of a batch: their streams already
    // overlap one another, and more streams than hardware queues would make
    // unrelated streams share a queue (and its order)
    hipStream_t cst = nullptr;
    hipEvent_t cev[2] = {nullptr, nullptr};
    // uploads of the next FASTQ block while the current one runs
    // (curc_fastq_stream_run); created on first use
    hipStream_t ust = nullptr;
    // pinned host buffers of host-parsed FASTQ sequences (tuning fq_host),
    // two so that one is filled while the other is copied; grown on demand
    uint8_t* hstage[2] = {nullptr, nullptr};
    uint64_t hstage_cap[2] = {0, 0};
    uint8_t* stage(int k, uint64_t bytes) {
        if (hstage_cap[k] < bytes) {
            if (hstage[k]) CURC_HIP(hipHostFree(hstage[k]));
            hstage[k] = nullptr;
            hstage_cap[k] = 0;
            CURC_HIP(hipHostMalloc((void**)&hstage[k], bytes, hipHostMallocDefault));
            hstage_cap[k] = bytes;
        }
        return hstage[k];
    }
    // overlap probes and bucket searches of the last block (copied from the
    // device after the loop; outlives the block's stack frame)
    uint64_t ov_work[2] = {0, 0};
    bool side_copies = true;
    // context (own stream, arena and counters) that runs the N-read stage
    // beside the LQ stage of a single block; created on first use
    curc_ctx* helper = nullptr;
    hipEvent_t nev = nullptr;  // end of mapping on st, for the helper to wait on
    // start and end of the N-reads' mapping launch on the helper's stream
    hipEvent_t mev[2] = {nullptr, nullptr};
    Scratch S;
    std::string err;
    bool timing = false;
    std::vector<curc_ctx*> workers;  // extra stream/arena pairs of curc_batch_run
    // hash index layout: CSR arrays only, no 64-byte bucket slots, for the
    // prefix and the mapping index (environment CURC_INDEX=csr at context
    // creation; the layout an index of more than 2^28 buckets uses)
    bool index_csr = false;
    // offset loops replayed as HIP graphs: environment CURC_GRAPHS=1 at
    // context creation turns this on (measured about 1% slower than plain
    // stream launches on the 20M-read block, so off by default)
    GraphCache graphs;
    // learned mapping (CURC_RUN_LEARNED): the scorer's device arrays
    bool has_scorer = false;
    MlpDev scorer_net{};
    float scorer_threshold = 0.f;           // reads that already have a mapping
    float scorer_threshold_unmapped = 0.f;  // reads without one
    const MlpDev& scorer_dev() const { return scorer_net; }
    // traced mapping (CURC_RUN_TRACE): sampling of reads and record capacity
    uint32_t trace_every = 16;
    uint64_t trace_cap = 1u << 22;
    std::vector<StageStat> stats;
    std::vector<PendingTimer> pending;
    std::vector<hipEvent_t> event_pool;

    int stat_id(const char* name) {
        for (size_t i = 0; i < stats.size(); ++i)
            if (stats[i].name == name) return (int)i;
        stats.push_back(StageStat{name});
        return (int)stats.size() - 1;
    }
    hipEvent_t ev() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e;
        CURC_HIP(hipEventCreate(&e));
        return e;
    }
    // Sums pending event pairs into stats (call after a stream sync).
    void harvest() {
        for (auto& p : pending) {
            float ms = 0;
            CURC_HIP(hipEventElapsedTime(&ms, p.a, p.b));
            stats[p.stat].ms += ms;
            stats[p.stat].launches += 1;
            event_pool.push_back(p.a);
            event_pool.push_back(p.b);
        }
        pending.clear();
    }
    void add_bytes(const char* name, double bytes, double items) {
        int s = stat_id(name);
        stats[s].alg_bytes += bytes;
        stats[s].items += items;
    }
};

namespace {

// Brackets a region of stream work with HIP events when timing is on.
struct Region {
    curc_ctx* c;
    int stat = -1;
    hipEvent_t a{}, b{};
    Region(curc_ctx* ctx, const char* name) : c(ctx) {
        if (!c->timing) return;
        stat = c->stat_id(name);
        a = c->ev();
        b = c->ev();
        CURC_HIP(hipEventRecord(a, c->st));
    }
    ~Region() {
        if (stat < 0) return;
        if (hipEventRecord(b, c->st) == hipSuccess) c->pending.push_back(PendingTimer{stat, a, b});
    }
};

template <typename T>
void fill(curc_ctx* c, T* p, uint64_t n, T v) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n)), dim3(256), 0, c->st, p, n, v);
}

void set_cnt(curc_ctx* c, uint32_t* p, uint32_t v) {
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, c->st, p, v);
}

uint32_t read_cnt(curc_ctx* c, const uint32_t* p) {
    uint32_t v = 0;
    CURC_HIP(hipMemcpyAsync(&v, p, 4, hipMemcpyDeviceToHost, c->st));
    CURC_HIP(stream_wait(c->st));
    return v;
}

// Two stable compactions of one list (length *d_n, host bound n): in one
// single-pass launch (k_compact2, look-back state in scratch slot S_LB) when
// both fit its tile limit, else one after the other.  With the single pass,
// its look-back error flag is copied to *lb_err on the stream: the caller
// checks it after its next wait on the stream.
template <class PA, class EA, class PB, class EB>
void compact_pair(curc_ctx* c, const uint32_t* d_n, uint64_t n, PA pa, EA ea, uint32_t* outa, PB pb, EB eb,
                  uint32_t* outb, uint32_t* lb_err) {
    Scratch& S = c->S;
    *lb_err = 0;
    if (2 * compact2_tiles(n) <= COMPACT2_MAX_TILES) {
        LookbackState* lb = S.as<LookbackState>(S_LB, 1);
        fill<uint32_t>(c, (uint32_t*)lb, 4 + compact2_words(n, n), 0u);
        compact2(d_n, pa, ea, outa, n, d_n, pb, eb, outb, n, lb, c->st);
        CURC_HIP(hipMemcpyAsync(lb_err, &lb->err, 4, hipMemcpyDeviceToHost, c->st));
    } else {
        compact(d_n, pa, ea, outa, S, S_PRIM, c->st, n);
        compact(d_n, pb, eb, outb, S, S_PRIM, c->st, n);
    }
}

template <typename V, typename T>
void d2h(curc_ctx* c, V& out, const T* src, uint64_t n) {
    static_assert(sizeof(typename V::value_type) == sizeof(T), "element size");
    out.resize(n);
    if (n) CURC_HIP(hipMemcpyAsync(out.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, c->st));
}

// d2h on the copy stream, after the compute stream's work so far (event k),
// when the destination is pinned; else on the compute stream.
void join_side(curc_ctx* c) {
    if (c->cst) CURC_HIP(stream_wait(c->cst));
}

template <typename V, typename T>
void d2h_side(curc_ctx* c, V& out, const T* src, uint64_t n, int k) {
    static_assert(sizeof(typename V::value_type) == sizeof(T), "element size");
    if (!c->side_copies || n * sizeof(T) < HostPool::POOL_MIN_BYTES) {
        d2h(c, out, src, n);
        return;
    }
    if (!c->cst && hipStreamCreateWithFlags(&c->cst, hipStreamNonBlocking) != hipSuccess) {
        c->cst = nullptr;
        c->side_copies = false;
        d2h(c, out, src, n);
        return;
    }
    out.resize(n);
    CURC_HIP(hipEventRecord(c->cev[k], c->st));
    CURC_HIP(hipStreamWaitEvent(c->cst, c->cev[k], 0));
    CURC_HIP(hipMemcpyAsync(out.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, c->cst));
}

template <typename T>
void h2d(curc_ctx* c, T* dst, const T* src, uint64_t n) {
    if (n) CURC_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, c->st));
}

// Records of one assembled pseudogenome, in emission order (device).
struct AsmOut {
    uint32_t* rid = nullptr;  // read id in the id space of the link arrays
    uint64_t* pos = nullptr;  // position in the pseudogenome
    uint32_t* nb = nullptr;   // bases the read contributes
    uint32_t n = 0;
    uint64_t len = 0;
};

}  // namespace

struct curc_block {
    uint32_t L = 0, W = 0;
    uint64_t n_reads = 0;
    hvec<uint64_t> packed;
    hvec<uint32_t> nonN, Nids, sorted;
    hvec<uint32_t> next_dedup, prev_dedup, next_ov, prev_ov, next_pr, prev_pr;
    hvec<uint16_t> off_ov;
    hvec<uint32_t> next, prev;
    hvec<uint16_t> offset;
    hvec<uint32_t> hq, unmatched;
    hvec<uint64_t> hq_pg2, lq_pg2;
    uint64_t hq_len = 0, lq_len = 0;
    std::string hq_text, lq_text;
    bool hq_text_ok = false, lq_text_ok = false;
    hvec<curc_alignment> hq_rec, map_rec, lq_rec;
    hvec<uint8_t> map_mis;
    hvec<uint32_t> unmapped, unmapped_N;
    hvec<uint32_t> lq_next, lq_prev;
    hvec<uint16_t> lq_off;
    hvec<char> Nascii;
    hvec<char> nlq_text;
    hvec<curc_alignment> nlq_rec;
    // pre-entropy streams (CURC_RUN_STREAMS)
    hvec<char> s_strand;
    hvec<uint16_t> s_read_off, s_lq_off, s_nlq_off, s_mis_off;
    hvec<uint8_t> s_mis_count, s_mis_base, s_mis_table;
    hvec<uint32_t> s_pe_order;
    hvec<uint64_t> s_id_pos, s_mis_off_start;
    hvec<curc_alignment> s_sorted;
    // traced mapping (CURC_RUN_TRACE): scorer features and labels
    hvec<float> trace_x, trace_y;
    // pseudogenome self-match (CURC_RUN_REFMATCH): parts HQ, LQ, N-read LQ
    struct RmPart {
        hvec<char> text, map_off, map_len;
        hvec<uint64_t> matches, found;  // (src, len, dest) triples, with KEEP_STAGES
    } rm[3];
};

namespace {

std::string decode_pg(const hvec<uint64_t>& pg2, uint64_t len) {
    static const char D[4] = {'A', 'C', 'G', 'T'};
    std::string s(len, 'A');
    for (uint64_t i = 0; i < len; ++i) s[i] = D[(pg2[i >> 5] >> (2 * (31 - (i & 31)))) & 3];
    return s;
}

// Builds the 2-bit pseudogenome on the device from assembled records over
// the packed reads R; returns the device array (ceil(len/32)+1 words, zero
// tail).
uint64_t* build_pg(curc_ctx* c, const uint64_t* R, uint32_t W, const AsmOut& a, int slot) {
    const uint64_t words = (a.len + 31) / 32 + 1;
    uint64_t* pg = c->S.as<uint64_t>(slot, words);
    fill<uint64_t>(c, pg, words, 0ull);
    if (a.n)
        hipLaunchKernelGGL(k_pg_write, dim3(grid_for(a.n)), dim3(256), 0, c->st, R, W, (const uint32_t*)a.rid,
                           (const uint64_t*)a.pos, (const uint32_t*)a.nb, (uint64_t)a.n, (unsigned long long*)pg);
    return pg;
}

// Assembly over the visit list V (n reads; ids in the space of the link
// arrays, which hold id_space entries).  Records land in the three scratch
// slots starting at out_slot.  See assembly.cuh.
AsmOut assemble(curc_ctx* c, const uint32_t* V, uint32_t n, uint64_t id_space, const uint32_t* next,
                const uint32_t* prev, const uint16_t* offset, uint32_t L, uint32_t* cnt, int out_slot) {
    AsmOut out;
    out.n = n;
    Scratch& S = c->S;
    hipStream_t st = c->st;
    out.rid = S.as<uint32_t>(out_slot, (uint64_t)n + 1);
    out.pos = S.as<uint64_t>(out_slot + 1, (uint64_t)n + 1);
    out.nb = S.as<uint32_t>(out_slot + 2, (uint64_t)n + 1);
    if (n == 0) return out;
    uint32_t* rank = S.as<uint32_t>(S_A_RANK, id_space + 1);
    uint32_t* nxtR = S.as<uint32_t>(S_A_NXT, n);
    uint32_t* prvR = S.as<uint32_t>(S_A_PRV, n);
    uint16_t* offR = S.as<uint16_t>(S_A_OFF, n);
    JumpState* js[2] = {S.as<JumpState>(S_A_ST0, n), S.as<JumpState>(S_A_ST1, n)};
    const uint32_t g = grid_for(n);
    hipLaunchKernelGGL(k_asm_rank, dim3(g), dim3(256), 0, st, V, (uint64_t)n, rank);
    hipLaunchKernelGGL(k_asm_links, dim3(g), dim3(256), 0, st, V, (uint64_t)n, next, prev, offset,
                       (const uint32_t*)rank, nxtR, prvR, offR);
    // list ranking by a ruling set (assembly.cuh): rulers, their walks, jump
    // rounds over the rulers (2^Kr > rulers >= ruler chain or cycle length),
    // then every node's state into js[0]
    int cur = 0;
    uint32_t* members = S.as<uint32_t>(S_A_MEMB, n);
    {
        uint32_t* rul = S.as<uint32_t>(S_A_RUL, n);
        uint32_t* ridx = S.as<uint32_t>(S_A_RIDX, n);
        uint32_t* rlist = S.as<uint32_t>(S_A_RLIST, n);
        JumpState* rs0 = S.as<JumpState>(S_A_RS0, n);
        fill<uint32_t>(c, rul, n, INV);
        set_cnt(c, cnt + C_TMP, n);
        compact(cnt + C_TMP, PredRuler{prvR}, EmitRuler{prvR, rlist, ridx, rs0}, cnt + C_ASM, S, S_PRIM, st, n);
        const uint32_t nr = read_cnt(c, cnt + C_ASM);
        JumpState* rs[2] = {rs0, S.as<JumpState>(S_A_RS1, nr)};
        const uint32_t gr = grid_for(nr);
        hipLaunchKernelGGL(k_asm_walk, dim3(gr), dim3(256), 0, st, (uint64_t)nr, (const uint32_t*)rlist,
                           (const uint32_t*)nxtR, (const uint16_t*)offR, L, (const uint32_t*)ridx, rul, js[1], rs0);
        uint32_t Kr = 1;
        while ((1ull << Kr) <= nr) ++Kr;
        int rc = 0;
        for (uint32_t r = 0; r < Kr; ++r) {
            hipLaunchKernelGGL(k_asm_round, dim3(gr), dim3(256), 0, st, (const uint32_t*)nullptr, (uint64_t)nr,
                               (const JumpState*)rs[rc], rs[1 - rc]);
            rc = 1 - rc;
        }
        set_cnt(c, cnt + C_ASM, 0);
        hipLaunchKernelGGL(k_asm_final, dim3(g), dim3(256), 0, st, (uint64_t)n, (const uint32_t*)rul,
                           (const JumpState*)js[1], (const JumpState*)rs[rc], js[0], members, cnt + C_ASM);
        const uint32_t nu = read_cnt(c, cnt + C_ASM);
        if (nu) {
            // cycles without a ruler: the rounds over their members (an even
            // number, so that the states end in js[0]; more rounds only widen
            // a cycle's window, which still wraps the cycle)
            const uint32_t gu = grid_for(nu);
            hipLaunchKernelGGL(k_asm_init, dim3(gu), dim3(256), 0, st, (const uint32_t*)members, (uint64_t)nu,
                               (const uint32_t*)nxtR, (const uint32_t*)prvR, (const uint16_t*)offR, L, js[0]);
            uint32_t Ku = 1;
            while ((1ull << Ku) <= nu) ++Ku;
            Ku += Ku & 1u;
            for (uint32_t r = 0; r < Ku; ++r) {
                hipLaunchKernelGGL(k_asm_round, dim3(gu), dim3(256), 0, st, (const uint32_t*)members, (uint64_t)nu,
                                   (const JumpState*)js[cur], js[1 - cur]);
                cur = 1 - cur;
            }
        }
    }
    uint32_t* ckey = S.as<uint32_t>(S_A_CKEY, n);
    uint64_t* clen = S.as<uint64_t>(S_A_CLEN, n);
    // the cycles' members (ckey / clen hold their cycle and its cut key
    // until k_asm_tails writes them at chain heads)
    uint32_t* cid = ckey;
    unsigned long long* best = reinterpret_cast<unsigned long long*>(clen);
    set_cnt(c, cnt + C_ASM, 0);
    hipLaunchKernelGGL(k_asm_cyc_list, dim3(g), dim3(256), 0, st, (uint64_t)n, (const JumpState*)js[cur], members,
                       cnt + C_ASM, cid, best);
    const uint32_t ncyc = read_cnt(c, cnt + C_ASM);
    if (ncyc) {
        // the other nodes are final in both buffers after this copy
        CURC_HIP(hipMemcpyAsync(js[1 - cur], js[cur], (uint64_t)n * sizeof(JumpState), hipMemcpyDeviceToDevice, st));
        const uint32_t gc = grid_for(ncyc);
        uint32_t Kc = 1;
        while ((1ull << Kc) <= ncyc) ++Kc;
        // each cycle ranked from its minimum-rank member, cut where the reference cuts it
        hipLaunchKernelGGL(k_asm_cyc_open, dim3(gc), dim3(256), 0, st, (const uint32_t*)members, (uint64_t)ncyc,
                           (const uint32_t*)cid, (const uint32_t*)prvR, js[cur]);
        for (uint32_t r = 0; r < Kc; ++r) {
            hipLaunchKernelGGL(k_asm_round, dim3(gc), dim3(256), 0, st, (const uint32_t*)members, (uint64_t)ncyc,
                               (const JumpState*)js[cur], js[1 - cur]);
            cur = 1 - cur;
        }
        hipLaunchKernelGGL(k_asm_cyc_best, dim3(gc), dim3(256), 0, st, (const uint32_t*)members, (uint64_t)ncyc,
                           (const uint32_t*)cid, (const JumpState*)js[cur], (const uint16_t*)offR, best);
        hipLaunchKernelGGL(k_asm_cyc_cut, dim3(gc), dim3(256), 0, st, (const uint32_t*)members, (uint64_t)ncyc,
                           (const uint32_t*)cid, (const JumpState*)js[cur], (const uint16_t*)offR,
                           (const unsigned long long*)best, nxtR, prvR);
        // the cut cycles ranked again as paths
        hipLaunchKernelGGL(k_asm_init, dim3(gc), dim3(256), 0, st, (const uint32_t*)members, (uint64_t)ncyc,
                           (const uint32_t*)nxtR, (const uint32_t*)prvR, (const uint16_t*)offR, L, js[cur]);
        for (uint32_t r = 0; r < Kc; ++r) {
            hipLaunchKernelGGL(k_asm_round, dim3(gc), dim3(256), 0, st, (const uint32_t*)members, (uint64_t)ncyc,
                               (const JumpState*)js[cur], js[1 - cur]);
            cur = 1 - cur;
        }
    }
    uint32_t* ccnt = S.as<uint32_t>(S_A_CCNT, n);
    hipLaunchKernelGGL(k_asm_tails, dim3(g), dim3(256), 0, st, (uint64_t)n, (const JumpState*)js[cur],
                       (const uint32_t*)nxtR, ckey, clen, ccnt);
    uint32_t* hk0 = S.as<uint32_t>(S_A_HK0, n);
    uint32_t* hk1 = S.as<uint32_t>(S_A_HK1, n);
    uint32_t* hv0 = S.as<uint32_t>(S_A_HV0, n);
    uint32_t* hv1 = S.as<uint32_t>(S_A_HV1, n);
    set_cnt(c, cnt + C_TMP, n);
    compact(cnt + C_TMP, PredHead{prvR}, EmitHead{ckey, hk0, hv0}, cnt + C_ASM, S, S_PRIM, st, n);
    const uint32_t nh = read_cnt(c, cnt + C_ASM);
    radix_sort_pairs<uint32_t>(hk0, hv0, hk1, hv1, nh, 0, log2u((uint64_t)n + 1), S, S_PRIM, st);
    uint64_t* olen = S.as<uint64_t>(S_A_OLEN, (uint64_t)nh + 1);
    uint32_t* ocnt = S.as<uint32_t>(S_A_OCNT, (uint64_t)nh + 1);
    fill<uint64_t>(c, olen + nh, 1, 0ull);
    fill<uint32_t>(c, ocnt + nh, 1, 0u);
    hipLaunchKernelGGL(k_asm_gather_sizes, dim3(grid_for(nh)), dim3(256), 0, st, (const uint32_t*)hv0, (uint64_t)nh,
                       (const uint64_t*)clen, (const uint32_t*)ccnt, olen, ocnt);
    scan_excl_t<uint64_t>(olen, olen, (uint64_t)nh + 1, S, S_PRIM, st);
    scan_excl_t<uint32_t>(ocnt, ocnt, (uint64_t)nh + 1, S, S_PRIM, st);
    uint64_t* cpos = S.as<uint64_t>(S_A_CPOS, n);
    uint32_t* crec = S.as<uint32_t>(S_A_CREC, n);
    hipLaunchKernelGGL(k_asm_scatter_bases, dim3(grid_for(nh)), dim3(256), 0, st, (const uint32_t*)hv0, (uint64_t)nh,
                       (const uint64_t*)olen, (const uint32_t*)ocnt, cpos, crec);
    hipLaunchKernelGGL(k_asm_records, dim3(g), dim3(256), 0, st, V, (uint64_t)n, (const JumpState*)js[cur],
                       (const uint32_t*)nxtR, (const uint16_t*)offR, L, (const uint64_t*)cpos,
                       (const uint32_t*)crec, out.rid, out.pos, out.nb);
    CURC_HIP(hipMemcpyAsync(&out.len, olen + nh, 8, hipMemcpyDeviceToHost, st));
    CURC_HIP(stream_wait(st));
    return out;
}

// Records of an assembly in host memory (word = pos << 2 | flag; read ids
// through the device array id_map when given).
void records_to_host(curc_ctx* c, const AsmOut& a, const uint32_t* id_map, uint64_t flag, hvec<curc_alignment>& rec) {
    curc_alignment* d = c->S.as<curc_alignment>(S_RECAOS, (uint64_t)a.n + 1);
    if (a.n)
        hipLaunchKernelGGL(k_rec_aos, dim3(grid_for(a.n)), dim3(256), 0, c->st, (const uint32_t*)a.rid,
                           (const uint64_t*)a.pos, id_map, (uint64_t)a.n, 2u, flag, d);
    d2h(c, rec, (const curc_alignment*)d, a.n);
}

// Exact-matching verdicts (k_ov_verdicts with kip 0, k 32) for reads of W
// words: its register copy of a read is sized at compile time.
inline void fm_verdicts(curc_ctx* c, const uint64_t* R, uint32_t W, uint32_t L, const uint32_t* sl,
                        const uint32_t* d_ns, uint32_t bound, const uint32_t* filt, uint32_t bmask, uint64_t* vb,
                        uint64_t vnw) {
    const dim3 vg(grid_for(bound, 65536));
#define CURC_FM_VERDICTS(WM)                                                                                     \
    hipLaunchKernelGGL((k_ov_verdicts<WM, true>), vg, dim3(256), 0, c->st, R, sl, bound, W, 32u, 0u, L - 31, filt, \
                       bmask, vb, vnw, d_ns)
    if (W <= 2) CURC_FM_VERDICTS(2);
    else if (W <= 3) CURC_FM_VERDICTS(3);
    else if (W <= 4) CURC_FM_VERDICTS(4);
    else if (W <= 5) CURC_FM_VERDICTS(5);
    else if (W <= 6) CURC_FM_VERDICTS(6);
    else if (W <= 7) CURC_FM_VERDICTS(7);
    else CURC_FM_VERDICTS(8);
#undef CURC_FM_VERDICTS
}

// Offsets [1, Ls) of the exact-matching loop over lists kept in place
// (k_fm_probe_s / k_fm_resolve_s): two launches per offset, and the block
// ranges of the prefix list once per block width.  For 2-bit reads of up to
// 8 words the suffixes are screened by the prefix filter at every offset
// that leaves at least 32 bases.  Links set the flags slinked / plinked, so
// that the lists can be compacted afterwards; the work counts of the phase
// are added to work[0] (suffix-offset pairs) and work[2] (searches), and
// work[1] gets the suffix count.
template <class E>
void fm_static_phase(curc_ctx* c, const uint64_t* R, uint32_t W, uint32_t L, uint32_t Ls, uint32_t n_max,
                     bool ranged, uint32_t MAX_BP, uint32_t BITS, uint32_t* cnt, uint32_t* next, uint32_t* prev,
                     uint16_t* offset, uint32_t* pl, uint64_t* pk, uint32_t* sl, uint32_t* d_np, uint32_t* d_ns,
                     uint32_t* choice, uint32_t* grp, uint32_t* memb, uint32_t* resl, uint32_t* pool,
                     uint64_t* kpool, uint8_t* slinked, uint8_t* plinked, uint8_t* lead, uint32_t* gcnt,
                     uint32_t* gmem, uint32_t* head, uint64_t* work, uint32_t g) {
    Scratch& S = c->S;
    hipStream_t st = c->st;
    uint32_t* rng = S.as<uint32_t>(S_RNG, (1u << 20) + 1);
    uint32_t* nf = S.as<uint32_t>(S_FMNF, n_max);
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n_max)), dim3(256), 0, st, nf, (uint64_t)n_max);
    fill<uint8_t>(c, slinked, n_max, (uint8_t)0);
    fill<uint8_t>(c, plinked, n_max, (uint8_t)0);
    // prefix filter and verdicts (offsets 1 .. L-32)
    const bool verd = E::BPW == 32 && W <= 8 && L > 32;
    uint64_t* vb = nullptr;
    const uint64_t vnw = ((uint64_t)n_max + 63) / 64;
    const uint32_t vmax = verd ? L - 31 : 1;  // verdicts for off < vmax
    if (verd) {
        uint64_t fbits = 1ull << 20;
        while (fbits < (uint64_t)c->tune.filter_bits * n_max && fbits < (1ull << 32)) fbits <<= 1;
        const uint32_t bmask = (uint32_t)(fbits / (32 * OVB_WORDS) - 1);
        uint32_t* filt = S.as<uint32_t>(S_FMFILT, fbits / 32);
        fill<uint32_t>(c, filt, fbits / 32, 0u);
        hipLaunchKernelGGL(k_fm_prefix_filter, dim3(grid_for(n_max)), dim3(256), 0, st, R, (const uint32_t*)pl,
                           (const uint32_t*)d_np, W, filt, bmask);
        vb = S.as<uint64_t>(S_FMVB, (uint64_t)(vmax - 1) * vnw);
        fm_verdicts(c, R, W, L, sl, d_ns, n_max, filt, bmask, vb, vnw);
    }
    uint32_t last_bp = 0;
    for (uint32_t off = 1; off < Ls; ++off) {
        const uint32_t bp = (L - off >= MAX_BP) ? MAX_BP : (L - off);
        if (ranged && bp != last_bp) {
            hipLaunchKernelGGL(k_fm_ranges<E>, dim3(grid_for((uint64_t)n_max + 1, 4096)), dim3(256), 0, st,
                               (const uint64_t*)pk, (const uint32_t*)d_np, bp, 1u << (BITS * bp), rng,
                               (uint8_t*)nullptr);
            last_bp = bp;
        }
        ClaimState cs{R, W, L, sl, d_ns, pl, pk, d_np, next, prev, offset, choice, grp, nullptr, nullptr,
                      head, memb, resl, cnt + C_RESN, pool, cnt + C_POOL, kpool, slinked, plinked, work, nf};
        hipLaunchKernelGGL(k_fm_probe_s<E>, dim3(g), dim3(256), 0, st, cs, ranged ? (const uint32_t*)rng : nullptr, bp,
                           off, (vb && off < vmax) ? (const uint64_t*)vb : nullptr, vnw, lead, gcnt, gmem, 0ull,
                           (uint32_t*)nullptr, 0ull);
        hipLaunchKernelGGL((k_fm_resolve_s<E>), dim3(3 * c->cus), dim3(256), 0, st, cs, off, lead, gcnt,
                           (const uint32_t*)gmem, 0ull);
    }
    hipLaunchKernelGGL(k_fm_count_work, dim3(grid_for(n_max, 1024)), dim3(256), 0, st, (const uint32_t*)sl,
                       (const uint32_t*)d_ns, (const uint32_t*)next, (const uint16_t*)offset, Ls,
                       (const uint64_t*)vb, vnw, vmax, work);
}

// Exact all-offset matching (a8/a11 over 2-bit reads, compress.cu:1003-1174;
// a12 over 3-bit N-reads, compress.cu:1176-1347).  Initial prefix/suffix
// lists are lexicographic; their lengths are the device counters
// cnt[C_PRE]/cnt[C_SUF], both <= n_max.  The block width bp only partitions
// the search; it does not change which prefixes a suffix meets.
template <class E>
void full_match(curc_ctx* c, const uint64_t* R, uint32_t W, uint32_t L, const uint32_t* d_pre,
                const uint32_t* d_suf, uint32_t* cnt, uint32_t n_max, uint32_t* next, uint32_t* prev,
                uint16_t* offset, const char* stage) {
    constexpr uint32_t BITS = (E::BPW == 32) ? 2 : 3;
    // block width: about one block per list entry, at most 2^20 block values;
    // short lists (<= FM_SHORT) go without block ranges, searching the whole
    // prefix list, which saves the fill / histogram / scan launches
    constexpr uint32_t FM_SHORT = 1u << 12;
    // lists up to this bound run the fused single-workgroup loop
    const uint32_t FM_FUSED = c->tune.fm_fused;
    if (n_max == 0) return;
    // bounds are rounded up to a power of two (>= FM_SHORT) so that blocks
    // of similar size share buffer sizes, grids and hence captured graphs;
    // the kernels read the true list lengths from device counters
    const uint32_t n_in = n_max;  // length of the caller's lists d_pre / d_suf
    {
        uint32_t b = FM_SHORT;
        while (b < n_max) b <<= 1;
        n_max = b;
    }
    uint32_t MAX_BP = 1;
    while (MAX_BP < ((BITS == 2) ? 10u : 6u) && (1ull << (BITS * MAX_BP)) < n_max) ++MAX_BP;
    const bool ranged = n_max > FM_SHORT;
    Scratch& S = c->S;
    uint32_t* pl[2] = {S.as<uint32_t>(S_PL0, n_max), S.as<uint32_t>(S_PL1, n_max)};
    uint64_t* pk[2] = {S.as<uint64_t>(S_PK0, n_max), S.as<uint64_t>(S_PK1, n_max)};
    uint32_t* sl[2] = {S.as<uint32_t>(S_SL0, n_max), S.as<uint32_t>(S_SL1, n_max)};
    uint32_t* choice = S.as<uint32_t>(S_CHOICE, n_max);
    uint32_t* grp = S.as<uint32_t>(S_GRP, n_max);
    uint32_t* memb = S.as<uint32_t>(S_MEMB, n_max);
    uint32_t* resl = S.as<uint32_t>(S_RESL, n_max);
    uint32_t* pool = S.as<uint32_t>(S_POOL, n_max);
    uint64_t* kpool = S.as<uint64_t>(S_KPOOL, n_max);
    uint8_t* slinked = S.as<uint8_t>(S_SLINK, n_max);
    uint8_t* plinked = S.as<uint8_t>(S_PLINK, n_max);
    uint8_t* lead = S.as<uint8_t>(S_FMLEAD, n_max);
    fill<uint8_t>(c, lead, n_max, (uint8_t)0);
    // per group (prefix position): member count and the first FM_SMALL members
    uint32_t* gcnt = S.as<uint32_t>(S_FMGC, n_max);
    uint32_t* gmem = S.as<uint32_t>(S_FMGM, (uint64_t)n_max * FM_SMALL);
    // over the compacted lists of the last offsets, rows widen as the lists shrink
    const uint64_t gcap = (uint64_t)n_max * FM_SMALL;
    fill<uint32_t>(c, gcnt, n_max, 0u);
    fill<uint32_t>(c, grp, n_max, INV);  // group marks (see k_fm_resolve_s)
    uint32_t* owner = S.as<uint32_t>(S_OWNER, (uint64_t)n_max + 1);
    uint32_t* cont = S.as<uint32_t>(S_CONT, (uint64_t)n_max + 1);
    uint32_t* head = S.as<uint32_t>(S_HEAD, (uint64_t)n_max + 1);
    uint32_t* rng = S.as<uint32_t>(S_RNG, (1u << 20) + 1);
    fill<uint32_t>(c, owner, (uint64_t)n_max + 1, INV);
    fill<uint32_t>(c, head, (uint64_t)n_max + 1, INV);
    fill<uint32_t>(c, cont, (uint64_t)n_max + 1, 0u);
    // work counts (see the byte count at the end)
    uint64_t* work = S.as<uint64_t>(S_WORK, 8);
    fill<uint64_t>(c, work, 8, 0ull);
    uint32_t* np[2] = {cnt + C_NP0, cnt + C_NP1};
    uint32_t* ns[2] = {cnt + C_NS0, cnt + C_NS1};
    CURC_HIP(hipMemcpyAsync(np[0], cnt + C_PRE, 4, hipMemcpyDeviceToDevice, c->st));
    CURC_HIP(hipMemcpyAsync(ns[0], cnt + C_SUF, 4, hipMemcpyDeviceToDevice, c->st));
    CURC_HIP(hipMemcpyAsync(pl[0], d_pre, (uint64_t)n_in * 4, hipMemcpyDeviceToDevice, c->st));
    CURC_HIP(hipMemcpyAsync(sl[0], d_suf, (uint64_t)n_in * 4, hipMemcpyDeviceToDevice, c->st));
    hipLaunchKernelGGL(k_gather_key0_dev, dim3(grid_for(n_max)), dim3(256), 0, c->st, R, (const uint32_t*)pl[0],
                       (const uint32_t*)np[0], W, pk[0]);
    const uint32_t g = grid_for(n_max, c->tune.fm_grid);
    if (!ranged && n_max <= FM_FUSED) {
        // short lists: the whole offset loop in one single-workgroup launch
        ClaimState c0{R, W, L, sl[0], ns[0], pl[0], pk[0], np[0], next, prev, offset,
                      choice, grp, owner, cont, head, memb, resl, cnt + C_RESN, pool, cnt + C_POOL,
                      kpool, slinked, plinked, work};
        ClaimState c1 = c0;
        c1.sl = sl[1];
        c1.d_ns = ns[1];
        c1.pl = pl[1];
        c1.pk = pk[1];
        c1.d_np = np[1];
        hipLaunchKernelGGL((k_fm_fused<E>), dim3(1), dim3(1024), 0, c->st, c0, c1, L);
        return;
    }
    // Offsets before the last FM_TAIL run over the lists in place (few reads
    // link per offset there); the lists are then compacted once, and the
    // tail, where most of the remaining reads link in large groups, runs
    // with both lists compacted after every offset.
    constexpr uint32_t FM_TAIL = 16;
    const uint32_t Ls = (c->tune.fm_static && L > FM_TAIL + 1) ? L - FM_TAIL : 1;
    // both lists compacted in one single-pass launch (up to COMPACT2_MAX_TILES
    // tiles), else by two three-launch compactions
    // (its status words are cleared by each offset's finalisation)
    LookbackState* lb = nullptr;
    uint32_t* lb_status = nullptr;
    uint64_t lb_words = 0;
    if (2 * compact2_tiles(n_max) <= COMPACT2_MAX_TILES) {
        lb = S.as<LookbackState>(S_LB, 1);
        fill<uint32_t>(c, (uint32_t*)lb, sizeof(LookbackState) / 4, 0u);
        lb_status = (uint32_t*)lb->status;
        lb_words = compact2_words(n_max, n_max);
    }
    auto compact_lists = [&](int cur) {
        if (lb) {
            compact2(ns[cur], PredFlagClear{slinked}, EmitGather{sl[cur], sl[1 - cur]}, ns[1 - cur], n_max, np[cur],
                     PredFlagClear{plinked}, EmitPre{pl[cur], pk[cur], pl[1 - cur], pk[1 - cur]}, np[1 - cur],
                     n_max, lb, c->st);
        } else {
            compact(ns[cur], PredFlagClear{slinked}, EmitGather{sl[cur], sl[1 - cur]}, ns[1 - cur], S, S_PRIM,
                    c->st, n_max);
            compact(np[cur], PredFlagClear{plinked}, EmitPre{pl[cur], pk[cur], pl[1 - cur], pk[1 - cur]},
                    np[1 - cur], S, S_PRIM, c->st, n_max);
        }
    };
    // the compacted tail's counts go to work[4] (suffix-list lengths) and
    // work[5] (the first offset's length when there is no in-place phase)
    const int cur0 = Ls > 1 ? 1 : 0;
    auto loop = [&]() {
        if (Ls > 1) {
            fm_static_phase<E>(c, R, W, L, Ls, n_max, ranged, MAX_BP, BITS, cnt, next, prev, offset, pl[0], pk[0],
                               sl[0], np[0], ns[0], choice, grp, memb, resl, pool, kpool, slinked, plinked, lead,
                               gcnt, gmem, head, work, g);
            compact_lists(0);
        }
        int cur = cur0;
        for (uint32_t off = Ls; off < L; ++off) {
            const uint32_t bp = (L - off >= MAX_BP) ? MAX_BP : (L - off);
            const uint32_t nb = 1u << (BITS * bp);
            if (ranged)
                hipLaunchKernelGGL(k_fm_ranges<E>, dim3(grid_for((uint64_t)n_max + 1, 4096)), dim3(256), 0, c->st,
                                   (const uint64_t*)pk[cur], (const uint32_t*)np[cur], bp, nb, rng, plinked);
            ClaimState cs{R, W, L, sl[cur], ns[cur], pl[cur], pk[cur], np[cur], next, prev, offset,
                          choice, grp, nullptr, nullptr, head, memb, resl, cnt + C_RESN, pool, cnt + C_POOL,
                          kpool, slinked, plinked, work + 4};
            hipLaunchKernelGGL(k_fm_probe_s<E>, dim3(g), dim3(256), 0, c->st, cs,
                               ranged ? (const uint32_t*)rng : nullptr, bp, off, (const uint64_t*)nullptr, 0ull, lead,
                               gcnt, gmem, gcap, lb_status, lb_words);
            // three workgroups per CU (the cooperative sort's LDS allows three)
            hipLaunchKernelGGL((k_fm_resolve_s<E>), dim3(3 * c->cus), dim3(256), 0, c->st, cs, off, lead, gcnt,
                               (const uint32_t*)gmem, gcap);
            compact_lists(cur);
            cur = 1 - cur;
        }
    };
    GraphKey key;
    key.v[0] = 1;  // the exact-matching loop
    key.v[1] = E::BPW;
    key.v[2] = W;
    key.v[3] = L;
    key.v[4] = n_max;
    key.v[5] = (uint64_t)R;
    key.v[6] = (uint64_t)next;
    key.v[7] = (uint64_t)prev;
    key.v[8] = (uint64_t)offset;
    key.v[9] = (uint64_t)cnt;
    key.v[10] = Ls;
    key.v[11] = g;
    key.v[12] = c->cus;
    c->graphs.run(key, S, c->st, loop);
    uint64_t w[8] = {};
    uint32_t fin = 0, lb_err = 0;
    CURC_HIP(hipMemcpyAsync(w, work, sizeof(w), hipMemcpyDeviceToHost, c->st));
    CURC_HIP(hipMemcpyAsync(&fin, ns[cur0 ^ ((L - Ls) & 1)], 4, hipMemcpyDeviceToHost, c->st));
    if (lb) CURC_HIP(hipMemcpyAsync(&lb_err, &lb->err, 4, hipMemcpyDeviceToHost, c->st));
    CURC_HIP(stream_wait(c->st));
    if (lb_err != 0) throw std::runtime_error("exact matching: list compaction look-back did not complete");
    // In place: per offset the suffix flags (a byte each) and, where there
    // are verdicts, their row (a bit each); per search the suffix's read and
    // block range.  Compacted: per offset and listed suffix its read and id.
    // Per link the prefix read.
    const double P8 = 8.0 * W;
    const uint64_t ns0 = std::max(w[1], w[5]);
    double bytes = (double)w[4] * (P8 + 8) + (double)(ns0 - std::min<uint64_t>(ns0, fin)) * (4 + P8);
    if (Ls > 1) {
        const bool verd = E::BPW == 32 && W <= 8 && L > 32;
        const uint32_t vrows = verd ? std::min(L - 31, Ls) - 1 : 0;
        bytes += (double)(Ls - 1) * ns0 + (double)vrows * ns0 / 8.0 + (double)w[2] * (P8 + 8);
    }
    c->add_bytes(stage, bytes, (double)(w[0] + w[4]));
}

// Record sort and stream builders (a13; compress.cu:1730-2044), on the
// device where the records are.  Fills the s_* fields of B.
// ------------------------------------------------ pseudogenome self-match ---
// RefMatcher::init (compress.cu:698-741): CSR index of source positions 5i,
// the first 12 per bucket in descending order.
struct RmIndex {
    uint64_t count = 0;
    uint32_t H = 0;
    uint32_t* kr = nullptr;
    uint32_t* pos = nullptr;
};

RmIndex rm_index(curc_ctx* c, const uint64_t* src, uint64_t src_len) {
    Scratch& S = c->S;
    hipStream_t st = c->st;
    RmIndex ix;
    if (src_len < RM_K) return ix;
    ix.count = (src_len - RM_K + 1 + RM_SRC_STEP - 1) / RM_SRC_STEP;
    if (ix.count > (1ull << 30)) throw std::length_error("pseudogenome too long for the self-match index");
    ix.H = hash_size(ix.count);
    uint32_t* k0 = S.as<uint32_t>(S_RM_K0, ix.count + 1);
    uint32_t* k1 = S.as<uint32_t>(S_RM_K1, ix.count + 1);
    uint32_t* v0 = S.as<uint32_t>(S_RM_V0, ix.count + 1);
    uint32_t* v1 = S.as<uint32_t>(S_RM_V1, ix.count + 1);
    uint32_t* krf = S.as<uint32_t>(S_RM_KRF, (uint64_t)ix.H + 2);
    ix.kr = S.as<uint32_t>(S_RM_KR, (uint64_t)ix.H + 2);
    ix.pos = S.as<uint32_t>(S_RM_POS, ix.count + 2);
    hipLaunchKernelGGL(k_rm_hash, dim3(grid_for(ix.count)), dim3(256), 0, st, src, ix.count, ix.H - 1, k0, v0);
    radix_sort_pairs<uint32_t>(k0, v0, k1, v1, ix.count, 0, log2u(ix.H), S, S_PRIM, st);
    hipLaunchKernelGGL(k_gap_fill, dim3(grid_for(ix.count + 1)), dim3(256), 0, st, (const uint32_t*)k0, ix.count,
                       (uint64_t)ix.H + 1, krf);
    hipLaunchKernelGGL(k_map_caps, dim3(grid_for((uint64_t)ix.H + 2)), dim3(256), 0, st, (const uint32_t*)krf,
                       (uint64_t)ix.H, ix.kr);
    scan_excl(ix.kr, ix.kr, (uint64_t)ix.H + 2, S, S_PRIM, st);
    hipLaunchKernelGGL(k_map_fill, dim3(grid_for(ix.count)), dim3(256), 0, st, (const uint32_t*)k0,
                       (const uint32_t*)v0, ix.count, (const uint32_t*)krf, (const uint32_t*)ix.kr, ix.pos);
    c->add_bytes("refmatch_index", (double)ix.count * (16 + 8) + 8.0 * ix.H, (double)ix.count);
    return ix;
}

// k_map_both for the read's word count W <= 8 (register capacity 3, 4, 5 or
// 8) and the index form (a.cs: compact slots).
template <bool CS, bool NR>
void launch_map_both_w(uint32_t W, dim3 grid, hipStream_t st, const MapArgs& a) {
    if (W <= 3)
        hipLaunchKernelGGL((k_map_both<3, CS, NR>), grid, dim3(256), 0, st, a);
    else if (W <= 4)
        hipLaunchKernelGGL((k_map_both<4, CS, NR>), grid, dim3(256), 0, st, a);
    else if (W <= 5)
        hipLaunchKernelGGL((k_map_both<5, CS, NR>), grid, dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((k_map_both<8, CS, NR>), grid, dim3(256), 0, st, a);
}
// the other reads, then the N-reads (grid capped by cap)
template <bool CS>
// The N-reads' launch goes to stream st_n (st, or a stream that runs it
// beside the other reads' launch), first.
void launch_map_both_t(uint32_t W, uint32_t cap, hipStream_t st, const MapArgs& a, hipStream_t st_n) {
    if (a.n_N) launch_map_both_w<CS, true>(W, dim3(grid_for(a.n_N, cap)), st_n, a);
    if (a.n_un) launch_map_both_w<CS, false>(W, dim3(grid_for(a.n_un, cap)), st, a);
}
void launch_map_both(uint32_t W, uint32_t cap, hipStream_t st, const MapArgs& a, hipStream_t st_n) {
    if (a.cs)
        launch_map_both_t<true>(W, cap, st, a, st_n);
    else
        launch_map_both_t<false>(W, cap, st, a, st_n);
}

// k_map_fast for the read's word count W <= 8 (register capacity 3, 4, 5 or 8).
void launch_map_fast(bool rc, uint32_t W, dim3 grid, hipStream_t st, const MapArgs& a) {
#define CURC_MAP_FAST(WM)                                                            \
    do {                                                                             \
        if (rc)                                                                      \
            hipLaunchKernelGGL((k_map_fast<true, WM>), grid, dim3(256), 0, st, a);   \
        else                                                                         \
            hipLaunchKernelGGL((k_map_fast<false, WM>), grid, dim3(256), 0, st, a);  \
    } while (0)
    if (W <= 3)
        CURC_MAP_FAST(3);
    else if (W == 4)
        CURC_MAP_FAST(4);
    else if (W == 5)
        CURC_MAP_FAST(5);
    else
        CURC_MAP_FAST(8);
#undef CURC_MAP_FAST
}

// k_map_range in register form for W <= 8 with the bucket-inline index
// (wbk = W; 0 for the CSR index), on one strand's slots or the compact
// two-strand index (a.cs), else in generic form.
void launch_map_range(bool rc, uint32_t wbk, dim3 grid, hipStream_t st, const MapArgs& a, const MapState& ms,
                      const uint32_t* act, const uint32_t* nact, uint32_t t0, uint32_t t1, float* x) {
#define CURC_MAP_RANGE_CS(WM, CS)                                                                                 \
    do {                                                                                                          \
        if (rc)                                                                                                   \
            hipLaunchKernelGGL((k_map_range<true, WM, CS>), grid, dim3(256), 0, st, a, ms, act, nact, t0, t1, x);  \
        else                                                                                                      \
            hipLaunchKernelGGL((k_map_range<false, WM, CS>), grid, dim3(256), 0, st, a, ms, act, nact, t0, t1, x); \
    } while (0)
#define CURC_MAP_RANGE(WM)               \
    do {                                 \
        if (a.cs)                        \
            CURC_MAP_RANGE_CS(WM, true);  \
        else                             \
            CURC_MAP_RANGE_CS(WM, false); \
    } while (0)
    if (wbk == 0 || wbk > 8)
        CURC_MAP_RANGE_CS(0, false);
    else if (wbk <= 3)
        CURC_MAP_RANGE(3);
    else if (wbk == 4)
        CURC_MAP_RANGE(4);
    else if (wbk == 5)
        CURC_MAP_RANGE(5);
    else
        CURC_MAP_RANGE(8);
#undef CURC_MAP_RANGE
#undef CURC_MAP_RANGE_CS
}

// "key=value,key=value" over the Tuning fields (unknown keys ignored).
void parse_tuning(const char* text, Tuning& t) {
    std::string s(text);
    size_t i = 0;
    while (i < s.size()) {
        size_t j = s.find(',', i);
        if (j == std::string::npos) j = s.size();
        const std::string kv = s.substr(i, j - i);
        const size_t e = kv.find('=');
        if (e != std::string::npos) {
            const std::string k = kv.substr(0, e);
            const uint32_t v = (uint32_t)std::strtoul(kv.c_str() + e + 1, nullptr, 10);
            if (k == "fm_fused") t.fm_fused = v;
            if (k == "ov_queue" && v <= 3) t.ov_queue = v;
            if (k == "map_both" && v <= 2) t.map_both = v;
            if (k == "fm_static" && v <= 1) t.fm_static = v;
            if (k == "fq_host" && v <= CURC_FQ_HOST_MAX) t.fq_host = v;
            if (v > 0) {
                if (k == "ov_grid") t.ov_grid = v;
                else if (k == "fm_grid") t.fm_grid = v;
                else if (k == "map_grid") t.map_grid = v;
                else if (k == "radix_items" && (v == 8 || v == 16 || v == 32)) t.radix_items = v;
                else if (k == "filter_bits" && v >= 8 && v <= 128 && (v & (v - 1)) == 0) t.filter_bits = v;
                else if (k == "workers" && v <= CURC_BATCH_MAX_WORKERS) t.workers = v;
            }
        }
        i = j + 1;
    }
}

template <typename T>
T read_one(curc_ctx* c, const T* p) {
    T v{};
    CURC_HIP(hipMemcpyAsync(&v, p, sizeof(T), hipMemcpyDeviceToHost, c->st));
    CURC_HIP(stream_wait(c->st));
    return v;
}

void put_triples(curc_ctx* c, hvec<uint64_t>& out, const uint64_t* s, const uint64_t* l, const uint64_t* d,
                 uint64_t n) {
    std::vector<uint64_t> a, b, e;
    d2h(c, a, s, n);
    d2h(c, b, l, n);
    d2h(c, e, d, n);
    CURC_HIP(stream_wait(c->st));
    out.resize(3 * n);
    for (uint64_t i = 0; i < n; ++i) {
        out[3 * i] = a[i];
        out[3 * i + 1] = b[i];
        out[3 * i + 2] = e[i];
    }
}

// RefMatcher::match (compress.cu:855-994) of one destination text, given as
// 2-bit words (dest2) or ASCII (desta, the character path for N-reads).
void rm_match(curc_ctx* c, const RmIndex& ix, const uint64_t* src, uint64_t src_len, const uint64_t* dest2,
              const uint8_t* desta, uint64_t dest_len, bool is_ref, bool keep, uint32_t* cnt,
              curc_block::RmPart& out) {
    Scratch& S = c->S;
    hipStream_t st = c->st;
    const bool binary = desta == nullptr;
    uint64_t n_found = 0, n_acc = 0, a_lastend = 0;
    bool any = false;
    uint64_t *as_ = nullptr, *a_start = nullptr, *a_end = nullptr, *goff = nullptr;
    uint32_t* len_off = nullptr;
    if (dest_len >= RM_K && ix.count) {
        const uint64_t dic = (dest_len - RM_K + 1 + RM_DEST_STEP - 1) / RM_DEST_STEP;
        const uint64_t capacity = (uint64_t)((double)dic * (is_ref ? 0.2 : 0.21));  // compress.cu:681-682
        const uint64_t n_slots = binary ? (dic + RM_SEG - 1) / RM_SEG : dic;
        if (n_slots >= 0xFFFFFFFFull) throw std::length_error("self-match destination too long");
        RmArgs a{};
        a.src = src;
        a.src_len = src_len;
        a.kr = ix.kr;
        a.pos = ix.pos;
        a.mask = ix.H - 1;
        a.dest_len = dest_len;
        a.n_slots = n_slots;
        a.is_ref = is_ref;
        if (binary) {
            const uint64_t words = (dest_len + 31) / 32 + 1;
            uint64_t* rc2 = S.as<uint64_t>(S_RM_RC, words);
            hipLaunchKernelGGL(k_rm_rc2, dim3(grid_for(words)), dim3(256), 0, st, dest2, dest_len, rc2, words);
            a.rc2 = rc2;
        } else {
            uint8_t* rca = S.as<uint8_t>(S_RM_RC, dest_len + 1);
            hipLaunchKernelGGL(k_rm_rc_ascii, dim3(grid_for(dest_len)), dim3(256), 0, st, desta, dest_len, rca);
            a.rca = rca;
        }
        a.r_src = S.as<uint64_t>(S_RM_S, n_slots);
        a.r_len = S.as<uint64_t>(S_RM_L, n_slots);
        a.r_dest = S.as<uint64_t>(S_RM_D, n_slots);
        a.flag = S.as<uint8_t>(S_RM_FLAG, n_slots);
        a.work = S.as<unsigned long long>(S_RM_WORK, 2);
        fill<unsigned long long>(c, a.work, 2, 0ull);
        if (binary)
            hipLaunchKernelGGL(k_rm_match2, dim3(grid_for(n_slots, 1u << 20)), dim3(256), 0, st, a);
        else
            hipLaunchKernelGGL(k_rm_match_ascii, dim3(grid_for(n_slots, 1u << 20)), dim3(256), 0, st, a);
        // results in collection order (compress.cu:794-797, 843-846)
        uint64_t* fs = S.as<uint64_t>(S_RM_FS, n_slots);
        uint64_t* fl = S.as<uint64_t>(S_RM_FL, n_slots);
        uint64_t* fd = S.as<uint64_t>(S_RM_FD, n_slots);
        set_cnt(c, cnt + C_TMP, (uint32_t)n_slots);
        compact(cnt + C_TMP, PredFlag8{a.flag}, EmitTriple{a.r_src, a.r_len, a.r_dest, fs, fl, fd}, cnt + C_RM, S,
                S_PRIM, st, n_slots);
        n_found = read_cnt(c, cnt + C_RM);
        if (c->timing) {
            unsigned long long wk[2];
            CURC_HIP(hipMemcpyAsync(wk, a.work, sizeof(wk), hipMemcpyDeviceToHost, st));
            CURC_HIP(stream_wait(st));
            // per probe: the 36-mer (9 B of 2-bit text) and two CSR offsets;
            // per verified candidate: its position and both 36-mers; per
            // extended base: a quarter byte of each text
            const double probes = binary ? (double)dic : (double)dic * 4.5;
            c->add_bytes("refmatch", probes * 17 + (double)wk[0] * 22 + (double)wk[1] * 0.5, (double)dic);
        }
        if (keep) put_triples(c, out.found, fs, fl, fd, n_found);
        any = n_found > 0;
        const uint64_t n = std::min(n_found, capacity);
        uint64_t n_u = 0;
        uint64_t *us = nullptr, *ul = nullptr, *ud = nullptr;
        if (n) {
            // compress.cu:914-932: transform, sort by (dest, src, len), unique
            uint64_t* k0 = S.as<uint64_t>(S_RM_SK0, n);
            uint64_t* k1 = S.as<uint64_t>(S_RM_SK1, n);
            uint32_t* v0 = S.as<uint32_t>(S_RM_SV0, n);
            uint32_t* v1 = S.as<uint32_t>(S_RM_SV1, n);
            hipLaunchKernelGGL(k_rm_transform, dim3(grid_for(n)), dim3(256), 0, st, fs, fl, fd, n, dest_len, is_ref,
                               k0, v0);
            radix_sort_pairs<uint64_t>(k0, v0, k1, v1, n, 0, log2u(dest_len + 1), S, S_PRIM, st);
            hipLaunchKernelGGL(k_rm_gather_key, dim3(grid_for(n)), dim3(256), 0, st, (const uint64_t*)fs,
                               (const uint32_t*)v0, n, k0);
            radix_sort_pairs<uint64_t>(k0, v0, k1, v1, n, 0, log2u(std::max(src_len, dest_len) + 1), S, S_PRIM, st);
            hipLaunchKernelGGL(k_rm_gather_key, dim3(grid_for(n)), dim3(256), 0, st, (const uint64_t*)fd,
                               (const uint32_t*)v0, n, k0);
            radix_sort_pairs<uint64_t>(k0, v0, k1, v1, n, 0, log2u(dest_len + 1), S, S_PRIM, st);
            us = S.as<uint64_t>(S_RM_US, n);
            ul = S.as<uint64_t>(S_RM_UL, n);
            ud = S.as<uint64_t>(S_RM_UD, n);
            set_cnt(c, cnt + C_TMP, (uint32_t)n);
            compact(cnt + C_TMP, PredUniqueTriple{fs, fl, fd, v0}, EmitTripleIdx{fs, fl, fd, v0, us, ul, ud},
                    cnt + C_RM, S, S_PRIM, st, n);
            n_u = read_cnt(c, cnt + C_RM);
            if (keep) put_triples(c, out.matches, us, ul, ud, n_u);
        }
        if (n_u) {
            // accepted matches (compress.cu:960-986) as a pointer chain
            const uint64_t ntile = (n_u + RM_TILE - 1) / RM_TILE;
            uint64_t* T = S.as<uint64_t>(S_RM_T, n_u);
            uint64_t* tmax = S.as<uint64_t>(S_RM_TMAX, ntile);
            hipLaunchKernelGGL(k_rm_reach, dim3((uint32_t)ntile), dim3(RM_TILE), 0, st, (const uint64_t*)ul,
                               (const uint64_t*)ud, n_u, T, tmax);
            uint32_t* J[2] = {S.as<uint32_t>(S_RM_J0, n_u + 2), S.as<uint32_t>(S_RM_J1, n_u + 2)};
            hipLaunchKernelGGL(k_rm_next, dim3(grid_for(n_u + 2)), dim3(256), 0, st, (const uint64_t*)ul,
                               (const uint64_t*)ud, (const uint64_t*)T, (const uint64_t*)tmax, n_u, J[0]);
            uint8_t* mark = S.as<uint8_t>(S_RM_MARK, n_u + 2);
            fill<uint8_t>(c, mark, n_u + 2, 0);
            fill<uint8_t>(c, mark + n_u, 1, 1);
            // after round r every chain node within 2^(r+1)-1 steps is marked
            uint32_t R = 0;
            while ((1ull << R) < n_u + 2) ++R;
            int cur = 0;
            for (uint32_t r = 0; r < R; ++r) {
                hipLaunchKernelGGL(k_rm_mark, dim3(grid_for(n_u + 2)), dim3(256), 0, st, (const uint32_t*)J[cur],
                                   n_u + 2, mark);
                if (r + 1 < R) {
                    hipLaunchKernelGGL(k_rm_jump, dim3(grid_for(n_u + 2)), dim3(256), 0, st, (const uint32_t*)J[cur],
                                       n_u + 2, J[1 - cur]);
                    cur = 1 - cur;
                }
            }
            as_ = S.as<uint64_t>(S_RM_AS, n_u);
            uint64_t* al = S.as<uint64_t>(S_RM_AL, n_u);
            uint64_t* ad = S.as<uint64_t>(S_RM_AD, n_u);
            set_cnt(c, cnt + C_TMP, (uint32_t)n_u);
            compact(cnt + C_TMP, PredFlag8{mark}, EmitTriple{us, ul, ud, as_, al, ad}, cnt + C_RM, S, S_PRIM, st,
                    n_u);
            n_acc = read_cnt(c, cnt + C_RM);
            a_start = S.as<uint64_t>(S_RM_ASTART, n_acc + 1);
            a_end = S.as<uint64_t>(S_RM_AEND, n_acc + 1);
            len_off = S.as<uint32_t>(S_RM_LB, n_acc + 1);
            goff = S.as<uint64_t>(S_RM_GAP, n_acc + 1);
            fill<uint32_t>(c, len_off + n_acc, 1, 0u);
            fill<uint64_t>(c, goff + n_acc, 1, 0ull);
            if (n_acc)
                hipLaunchKernelGGL(k_rm_accept, dim3(grid_for(n_acc)), dim3(256), 0, st, (const uint64_t*)as_,
                                   (const uint64_t*)al, (const uint64_t*)ad, n_acc, a_start, a_end, len_off, goff);
            scan_excl(len_off, len_off, n_acc + 1, S, S_PRIM, st);
            scan_excl_t<uint64_t>(goff, goff, n_acc + 1, S, S_PRIM, st);
            if (n_acc) a_lastend = read_one(c, a_end + n_acc - 1);
        }
    }
    // rewritten text (compress.cu:953-993; texts are the reverse complement
    // of a reverse complement, compress.cu:873 and 954)
    uint64_t gap_total = 0;
    if (goff) gap_total = read_one(c, goff + n_acc);
    const uint64_t out_len = gap_total + (dest_len - a_lastend);
    uint8_t* text = S.as<uint8_t>(S_RM_TEXT, out_len + 1);
    if (!goff) {
        goff = S.as<uint64_t>(S_RM_GAP, 1);
        fill<uint64_t>(c, goff, 1, 0ull);
    }
    if (out_len)
        hipLaunchKernelGGL(k_rm_text, dim3(grid_for((out_len + 63) / 64)), dim3(256), 0, st, dest2, desta,
                           (const uint64_t*)a_start, (const uint64_t*)a_end, (const uint64_t*)goff, n_acc, dest_len,
                           out_len, text);
    out.text.resize(out_len);
    if (out_len) CURC_HIP(hipMemcpyAsync(&out.text[0], text, out_len, hipMemcpyDeviceToHost, st));
    out.map_off.clear();
    out.map_len.clear();
    if (any) {
        const bool off32 = src_len <= 0xFFFFFFFFull;
        const uint32_t w = off32 ? 4 : 8;
        const uint32_t len_bytes = n_acc ? read_one(c, len_off + n_acc) : 0;
        uint8_t* moff = S.as<uint8_t>(S_RM_MOFF, n_acc * w + 1);
        uint8_t* mlen = S.as<uint8_t>(S_RM_MLEN, (uint64_t)len_bytes + 1);
        fill<uint8_t>(c, mlen, 1, (uint8_t)RM_TARGET);  // helper.cpp:179-187 of 50: one byte
        if (n_acc)
            hipLaunchKernelGGL(k_rm_maps, dim3(grid_for(n_acc)), dim3(256), 0, st, (const uint64_t*)as_,
                               (const uint64_t*)a_start, (const uint64_t*)a_end, (const uint32_t*)len_off, n_acc,
                               off32, moff, mlen);
        out.map_off.resize(n_acc * w);
        out.map_len.resize((uint64_t)len_bytes + 1);
        if (n_acc) CURC_HIP(hipMemcpyAsync(&out.map_off[0], moff, n_acc * w, hipMemcpyDeviceToHost, st));
        CURC_HIP(hipMemcpyAsync(&out.map_len[0], mlen, (uint64_t)len_bytes + 1, hipMemcpyDeviceToHost, st));
    }
    CURC_HIP(stream_wait(st));
}

// ------------------------------------------------- mapping in probe ranges ---
// Probe counts after which the scorer's checkpoints run.
constexpr uint32_t MAP_CHECKPOINTS[] = {2, 4, 8, 16, 32};

// One strand pass (one index part) of the mapping as probe ranges with the
// loop state in HBM.  Learned: at each checkpoint the network scores every
// active read and reads below the threshold end the pass.  Trace: sampled
// reads' features are recorded at each checkpoint and labelled at the end
// of the pass; no read is stopped, so the results equal k_map's.
void map_phased(curc_ctx* c, const MapArgs& a, bool rc, uint64_t nm, uint32_t* cnt, bool learned, bool trace) {
    Scratch& S = c->S;
    hipStream_t st = c->st;
    if (nm == 0) return;
    uint32_t* act[2] = {S.as<uint32_t>(S_MACT0, nm + 1), S.as<uint32_t>(S_MACT1, nm + 1)};
    uint32_t* nact[2] = {cnt + C_MA0, cnt + C_MA1};
    set_cnt(c, cnt + C_TMP, (uint32_t)nm);
    compact(cnt + C_TMP, PredNeedsMap{a.last}, EmitIota{act[0]}, nact[0], S, S_PRIM, st, nm);
    MapState ms{S.as<uint64_t>(S_MSEEN, 4 * nm), S.as<uint32_t>(S_MCNT, 4 * nm), S.as<uint8_t>(S_MSTOP, nm), nm};
    const uint32_t g = grid_for(nm, c->tune.map_grid);
    hipLaunchKernelGGL(k_map_state_reset, dim3(g), dim3(256), 0, st, a, ms, (const uint32_t*)act[0],
                       (const uint32_t*)nact[0]);
    float* X = (learned || trace) ? S.as<float>(S_MX, nm * MAP_NFEAT) : nullptr;
    float* logit = learned ? S.as<float>(S_MLOGIT, nm) : nullptr;
    unsigned long long* tcount = trace ? S.as<unsigned long long>(S_TRCNT, 1) : nullptr;
    uint64_t r0 = 0;
    if (trace) r0 = read_one(c, tcount);
    const uint32_t nprobe = (a.L - MAP_K) / MAP_READ_STEP + 1;
    int cur = 0;
    uint32_t t0 = 0;
    for (uint32_t ci = 0;; ++ci) {
        const uint32_t t1 = ci < sizeof(MAP_CHECKPOINTS) / 4 ? std::min(MAP_CHECKPOINTS[ci], nprobe) : nprobe;
        const bool last_range = t1 >= nprobe;
        // the range kernel writes each listed read's features at its end
        float* Xr = last_range ? nullptr : X;
        if (t1 > t0) launch_map_range(rc, a.bk ? a.W : 0, dim3(g), st, a, ms, act[cur], nact[cur], t0, t1, Xr);
        if (last_range) break;
        if (trace)
            hipLaunchKernelGGL(k_map_trace_emit, dim3(g), dim3(256), 0, st, (const uint32_t*)act[cur],
                               (const uint32_t*)nact[cur], (const float*)X, (const uint8_t*)a.last, c->trace_every,
                               S.as<float>(S_TRX, c->trace_cap * MAP_NFEAT), S.as<uint8_t>(S_TRLAST, c->trace_cap),
                               S.as<uint32_t>(S_TRSLOT, c->trace_cap), tcount, c->trace_cap);
        if (learned) {
            hipLaunchKernelGGL(k_mlp_infer, dim3(mlp_infer_grid(nm)), dim3(64 * MLP_INF_WAVES), 0, st,
                               c->scorer_dev(), (const float*)X, nm, (const uint32_t*)nact[cur], logit);
            hipLaunchKernelGGL(k_map_stop, dim3(g), dim3(256), 0, st, ms, (const uint32_t*)act[cur],
                               (const uint32_t*)nact[cur], (const float*)logit, (const uint8_t*)a.last,
                               a.max_mm, c->scorer_threshold, c->scorer_threshold_unmapped);
        }
        compact(nact[cur], PredMapActive{act[cur], a.last, ms.stop}, EmitGather{act[cur], act[1 - cur]},
                nact[1 - cur], S, S_PRIM, st, nm);
        cur = 1 - cur;
        t0 = t1;
    }
    if (trace) {
        const uint64_t r1 = std::min<uint64_t>(read_one(c, tcount), c->trace_cap);
        if (r1 > r0)
            hipLaunchKernelGGL(k_map_trace_label, dim3(grid_for(r1 - r0)), dim3(256), 0, st,
                               (const uint8_t*)S.as<uint8_t>(S_TRLAST, c->trace_cap),
                               (const uint32_t*)S.as<uint32_t>(S_TRSLOT, c->trace_cap), (const uint8_t*)a.last,
                               std::min(r0, c->trace_cap), r1, S.as<float>(S_TRY, c->trace_cap));
    }
}

void build_streams(curc_ctx* c, curc_block* B, const AsmOut& hq_asm, const AsmOut& lq_asm, const AsmOut& nlq_asm,
                   uint32_t* cnt, bool preserve, bool paired) {
    Scratch& S = c->S;
    hipStream_t st = c->st;
    const uint32_t L = B->L, W = B->W;
    curc_params P;
    curc_make_params(L, &P);
    const uint32_t max_mm = P.max_mm;
    const uint64_t N = B->n_reads;
    const uint32_t nmap = (uint32_t)B->map_rec.size();
    const uint32_t n = hq_asm.n + nmap;
    const uint32_t nN = (uint32_t)B->Nids.size();
    RecSrc src{hq_asm.rid, hq_asm.pos, hq_asm.n, S.as<uint32_t>(S_MREC_ID, (uint64_t)nmap + 1),
               S.as<uint64_t>(S_MREC_W, (uint64_t)nmap + 1), S.as<uint8_t>(S_MIS, (uint64_t)nmap + 1),
               S.as<uint32_t>(S_NIDS, (uint64_t)nN + 1)};
    // compress.cu:1743-1756 (stable; HQ records precede map records on ties)
    uint64_t* k0 = S.as<uint64_t>(S_K0, (uint64_t)n + 1);
    uint64_t* k1 = S.as<uint64_t>(S_K1, (uint64_t)n + 1);
    uint32_t* v0 = S.as<uint32_t>(S_V0, (uint64_t)n + 1);
    uint32_t* v1 = S.as<uint32_t>(S_V1, (uint64_t)n + 1);
    const uint32_t g = grid_for(n);
    if (n) hipLaunchKernelGGL(k_rec_keys, dim3(g), dim3(256), 0, st, src, n, preserve, k0, v0);
    const uint32_t kbits = preserve ? log2u(N + 1) : log2u(B->hq_len + 1);
    radix_sort_pairs<uint64_t>(k0, v0, k1, v1, n, 0, ((kbits + 7) / 8) * 8, S, S_PRIM, st);
    StreamOut o;
    o.rid = S.as<uint32_t>(S_ST_RID, (uint64_t)n + 1);
    o.rword = S.as<uint64_t>(S_ST_RWORD, (uint64_t)n + 1);
    o.strand = S.as<uint8_t>(S_ST_STRAND, (uint64_t)n + 1);
    o.mis_count = S.as<uint8_t>(S_ST_MISC, (uint64_t)n + 1);
    o.read_off = S.as<uint16_t>(S_ST_ROFF, (uint64_t)n + 1);
    const uint64_t n_pe = (uint64_t)n + lq_asm.n + nlq_asm.n;
    o.pe_order = S.as<uint32_t>(S_ST_PE, n_pe + 1);
    o.id_pos = S.as<uint64_t>(S_ST_IDPOS, N + 1);
    o.mkey = S.as<uint32_t>(S_ST_MKEY, (uint64_t)n + 1);
    o.midx = S.as<uint32_t>(S_ST_MIDX, (uint64_t)n + 1);
    if (preserve) fill<uint64_t>(c, o.id_pos, N, 0ull);
    if (n) hipLaunchKernelGGL(k_rec_streams, dim3(g), dim3(256), 0, st, src, (const uint32_t*)v0, n, preserve, paired, o);
    // mismatch-symbol index: exclusive scan of counts (total at [n])
    uint32_t* bidx = S.as<uint32_t>(S_ST_BIDX, (uint64_t)n + 1);
    CURC_HIP(hipMemcpyAsync(bidx, o.mkey, (uint64_t)n * 4, hipMemcpyDeviceToDevice, st));
    fill<uint32_t>(c, bidx + n, 1, 0u);
    scan_excl(bidx, bidx, (uint64_t)n + 1, S, S_PRIM, st);
    // offset-stream index: group records by mismatch count (stable)
    uint32_t* mk2 = S.as<uint32_t>(S_ST_MKEY2, (uint64_t)n + 1);
    uint32_t* mi2 = S.as<uint32_t>(S_ST_MIDX2, (uint64_t)n + 1);
    uint32_t* mk = o.mkey;
    uint32_t* mi = o.midx;
    radix_sort_pairs<uint32_t>(mk, mi, mk2, mi2, n, 0, 8, S, S_PRIM, st);
    uint32_t* gstart = S.as<uint32_t>(S_ST_GSTART, (uint64_t)max_mm + 3);
    hipLaunchKernelGGL(k_gap_fill, dim3(grid_for((uint64_t)n + 1)), dim3(256), 0, st, (const uint32_t*)mk, (uint64_t)n,
                       (uint64_t)max_mm + 1, gstart);
    std::vector<uint32_t> hgs(max_mm + 2);
    uint32_t total_mis = 0;
    CURC_HIP(hipMemcpyAsync(hgs.data(), gstart, (max_mm + 2) * 4, hipMemcpyDeviceToHost, st));
    CURC_HIP(hipMemcpyAsync(&total_mis, bidx + n, 4, hipMemcpyDeviceToHost, st));
    CURC_HIP(stream_wait(st));
    std::vector<uint64_t> mbase(max_mm + 2, 0);
    B->s_mis_off_start.assign(max_mm + 2, 0);
    uint64_t acc = 0;
    for (uint32_t m = 1; m <= max_mm; ++m) {
        mbase[m] = acc;
        B->s_mis_off_start[m] = acc;
        acc += (uint64_t)mis_width(m) * (hgs[m + 1] - hgs[m]);
    }
    B->s_mis_off_start[max_mm + 1] = acc;
    uint64_t* dmbase = S.as<uint64_t>(S_ST_MBASE, (uint64_t)max_mm + 2);
    h2d(c, dmbase, mbase.data(), max_mm + 2);
    uint32_t* oidx = S.as<uint32_t>(S_ST_OIDX, (uint64_t)n + 1);
    if (n)
        hipLaunchKernelGGL(k_rec_offidx, dim3(g), dim3(256), 0, st, (const uint32_t*)mk, (const uint32_t*)mi, n,
                           (const uint32_t*)gstart, (const uint64_t*)dmbase, oidx);
    // compress.cu:1815-1886: mismatch symbols and offsets
    uint8_t* pair = S.as<uint8_t>(S_ST_PAIR, (uint64_t)total_mis + 1);
    uint16_t* offs = S.as<uint16_t>(S_ST_OFFS, acc + 1);
    uint32_t* hist = S.as<uint32_t>(S_ST_HIST, 20);
    fill<uint32_t>(c, hist, 20, 0u);
    MisArgs ma{S.as<uint64_t>(S_PACKED, N * W + 1), W, L, S.as<uint8_t>(S_NASCII, (uint64_t)nN * L + 16),
               S.as<uint64_t>(S_PG_HQ, (B->hq_len + 31) / 32 + 1), o.rid, o.rword, o.mis_count, bidx, oidx,
               pair, offs, hist};
    if (n) hipLaunchKernelGGL(k_rec_mismatch, dim3(grid_for(n, 4096)), dim3(256), 0, st, ma, n);
    // compress.cu:1894-1929: per-context ranks by frequency, ties in
    // destination order (std::sort of four elements is an insertion sort)
    uint32_t hh[20];
    CURC_HIP(hipMemcpyAsync(hh, hist, sizeof(hh), hipMemcpyDeviceToHost, st));
    CURC_HIP(stream_wait(st));
    static const char CTX[4] = {'A', 'T', 'C', 'G'};
    static const char DEST[4][4] = {{'C', 'G', 'N', 'T'}, {'A', 'C', 'G', 'N'}, {'A', 'G', 'N', 'T'}, {'A', 'C', 'N', 'T'}};
    auto sym = [](char ch) { return ch == 'A' ? 0 : ch == 'C' ? 1 : ch == 'G' ? 2 : ch == 'T' ? 4 : 3; };
    uint8_t rank[20] = {0};
    B->s_mis_table.assign(16, 0);
    for (int cx = 0; cx < 4; ++cx) {
        int idx[4] = {0, 1, 2, 3};
        std::stable_sort(idx, idx + 4, [&](int a, int b) { return hh[cx * 5 + sym(DEST[cx][a])] > hh[cx * 5 + sym(DEST[cx][b])]; });
        for (int r = 0; r < 4; ++r) rank[cx * 5 + sym(DEST[cx][idx[r]])] = (uint8_t)r;
        for (int d = 0; d < 4; ++d) B->s_mis_table[cx * 4 + rank[cx * 5 + sym(DEST[cx][d])]] = (uint8_t)DEST[cx][d];
    }
    (void)CTX;
    uint8_t* drank = S.as<uint8_t>(S_ST_RANK, 32);
    h2d(c, drank, rank, 20);
    uint8_t* symb = S.as<uint8_t>(S_ST_SYM, (uint64_t)total_mis + 1);
    if (total_mis)
        hipLaunchKernelGGL(k_rec_recode, dim3(grid_for(total_mis)), dim3(256), 0, st, (const uint8_t*)pair,
                           (uint64_t)total_mis, (const uint8_t*)drank, symb);
    // compress.cu:1968-2001: LQ and N-LQ records
    uint16_t* lqoff = S.as<uint16_t>(S_ST_LQOFF, (uint64_t)lq_asm.n + 1);
    uint16_t* nlqoff = S.as<uint16_t>(S_ST_NLQOFF, (uint64_t)nlq_asm.n + 1);
    if (lq_asm.n)
        hipLaunchKernelGGL(k_lq_streams, dim3(grid_for(lq_asm.n)), dim3(256), 0, st, (const uint32_t*)lq_asm.rid,
                           (const uint64_t*)lq_asm.pos, lq_asm.n,
                           (const uint32_t*)S.as<uint32_t>(S_UNMAPPED, (uint64_t)lq_asm.n + 1), L, preserve, paired,
                           B->hq_len, lqoff, o.pe_order + n, o.id_pos);
    if (nlq_asm.n)
        hipLaunchKernelGGL(k_lq_streams, dim3(grid_for(nlq_asm.n)), dim3(256), 0, st, (const uint32_t*)nlq_asm.rid,
                           (const uint64_t*)nlq_asm.pos, nlq_asm.n, (const uint32_t*)src.Nids, L, preserve, paired,
                           B->hq_len + B->lq_len, nlqoff, o.pe_order + n + lq_asm.n, o.id_pos);
    // to host (the sorted records in their 16-byte layout, built on the device)
    curc_alignment* drec = S.as<curc_alignment>(S_RECAOS, (uint64_t)n + 1);
    if (n)
        hipLaunchKernelGGL(k_rec_aos, dim3(grid_for(n)), dim3(256), 0, st, (const uint32_t*)o.rid,
                           (const uint64_t*)o.rword, (const uint32_t*)nullptr, (uint64_t)n, 0u, 0ull, drec);
    d2h(c, B->s_sorted, (const curc_alignment*)drec, n);
    d2h(c, B->s_strand, (const char*)o.strand, n);
    d2h(c, B->s_mis_count, o.mis_count, n);
    d2h(c, B->s_mis_base, symb, total_mis);
    d2h(c, B->s_mis_off, offs, acc);
    if (!preserve) {
        d2h(c, B->s_read_off, o.read_off, n);
        d2h(c, B->s_lq_off, lqoff, lq_asm.n);
        d2h(c, B->s_nlq_off, nlqoff, nlq_asm.n);
        if (paired) d2h(c, B->s_pe_order, o.pe_order, n_pe);
    } else {
        d2h(c, B->s_id_pos, o.id_pos, N);
    }
    CURC_HIP(stream_wait(st));
}

// a12: N-read LQ pseudogenome (compress.cu:1724-1728, 1176-1347).  Same
// exact matching over the N-reads in the 3-bit ACGNT encoding; ids are N
// indices, and the visit order is the sorted unmapped list.  Runs on context
// x (its arena, counters and stream): the block's own, or a helper that then
// first waits for `wait` (the end of mapping on the block's stream).  Nascii
// and umN are the N-reads' ASCII and their unmapped list in device memory.
struct NlqOut {
    AsmOut a;
    uint8_t* npg = nullptr;
};
NlqOut nlq_stage(curc_ctx* x, hipEvent_t wait, const uint8_t* Nascii, uint64_t nN, const uint32_t* umN, uint32_t n_nlq,
                 uint32_t L, curc_block* B) {
    // ---- a12: N-read LQ pseudogenome (compress.cu:1724-1728, 1176-1347).
    // Same exact matching over the N-reads in the 3-bit ACGNT encoding; ids
    // are N indices, and the visit order is the sorted unmapped list.
    NlqOut out;
    Scratch& S = x->S;
    hipStream_t st = x->st;
    if (wait) CURC_HIP(hipStreamWaitEvent(st, wait, 0));
    uint32_t* cnt = S.as<uint32_t>(S_CNT, C_COUNT);
    if (wait) fill<uint32_t>(x, cnt, C_COUNT, 0u);  // a helper's own counters
    {
        Region r(x, "nlq");
        const uint32_t W3 = (L + 20) / 21;
        uint64_t* p3 = S.as<uint64_t>(S_N3, (uint64_t)nN * W3 + 1);
        hipLaunchKernelGGL(k_pack3, dim3(grid_for((uint64_t)nN * W3)), dim3(256), 0, st, (const uint8_t*)Nascii,
                           (uint64_t)nN, L, W3, p3);
        fill<uint64_t>(x, p3 + (uint64_t)nN * W3, 1, 0ull);
        uint32_t* nnext = S.as<uint32_t>(S_NNEXT, nN);
        uint32_t* nprev = S.as<uint32_t>(S_NPREV, nN);
        uint16_t* noff = S.as<uint16_t>(S_NOFF, nN);
        fill<uint32_t>(x, nnext, nN, INV);
        fill<uint32_t>(x, nprev, nN, INV);
        fill<uint16_t>(x, noff, nN, INVOFF);
        // compress.cu:1179-1193: sort by content (ties by N index), link equal reads
        uint64_t* k0 = S.as<uint64_t>(S_K0, (uint64_t)n_nlq + 1);
        uint64_t* k1 = S.as<uint64_t>(S_K1, (uint64_t)n_nlq + 1);
        uint32_t* v0 = S.as<uint32_t>(S_V0, (uint64_t)n_nlq + 1);
        uint32_t* v1 = S.as<uint32_t>(S_V1, (uint64_t)n_nlq + 1);
        CURC_HIP(hipMemcpyAsync(v0, umN, (uint64_t)n_nlq * 4, hipMemcpyDeviceToDevice, st));
        for (int w = (int)W3 - 1; w >= 0; --w) {
            uint32_t nd = (w == (int)W3 - 1) ? (L - 21 * (W3 - 1)) : 21;
            uint32_t begin = ((63 - 3 * nd) / 8) * 8;
            hipLaunchKernelGGL(k_gather_word, dim3(grid_for(n_nlq)), dim3(256), 0, st, (const uint64_t*)p3,
                               (const uint32_t*)v0, (uint64_t)n_nlq, W3, (uint32_t)w, k0);
            radix_sort_pairs<uint64_t>(k0, v0, k1, v1, n_nlq, begin, 64, S, S_PRIM, st);
        }
        uint32_t* nsorted = S.as<uint32_t>(S_NSORT, n_nlq);
        CURC_HIP(hipMemcpyAsync(nsorted, v0, (uint64_t)n_nlq * 4, hipMemcpyDeviceToDevice, st));
        if (n_nlq > 1)
            hipLaunchKernelGGL(k_dedup_link, dim3(grid_for(n_nlq)), dim3(256), 0, st, (const uint64_t*)p3,
                               (const uint32_t*)nsorted, (uint64_t)n_nlq, W3, nnext, nprev, noff,
                               S.as<uint8_t>(S_DEQ, (uint64_t)n_nlq + 1));
        uint32_t* pre = S.as<uint32_t>(S_PRE, n_nlq);
        uint32_t* suf = S.as<uint32_t>(S_SUF, n_nlq);
        set_cnt(x, cnt + C_TMP, n_nlq);
        uint32_t lb_err = 0;
        compact_pair(x, cnt + C_TMP, n_nlq, PredLinkState{nsorted, nprev}, EmitGather{nsorted, pre}, cnt + C_PRE,
                     PredLinkState{nsorted, nnext}, EmitGather{nsorted, suf}, cnt + C_SUF, &lb_err);
        CURC_HIP(stream_wait(st));
        if (lb_err != 0) throw std::runtime_error("N-read list compaction: look-back did not complete");
        full_match<Enc3>(x, p3, W3, L, pre, suf, cnt, n_nlq, nnext, nprev, noff, "nlq");
        out.a = assemble(x, nsorted, n_nlq, nN, nnext, nprev, noff, L, cnt, S_NLQIDS);
        const AsmOut& a = out.a;
        uint8_t* npg = S.as<uint8_t>(S_NPG, a.len + 16);
        out.npg = npg;
        if (a.n)
            hipLaunchKernelGGL(k_pg_write_ascii, dim3(grid_for(a.n)), dim3(256), 0, st, (const uint8_t*)Nascii, L,
                               (const uint32_t*)a.rid, (const uint64_t*)a.pos, (const uint32_t*)a.nb,
                               (uint64_t)a.n, npg);
        B->nlq_text.resize(a.len);
        if (a.len) CURC_HIP(hipMemcpyAsync(&B->nlq_text[0], npg, a.len, hipMemcpyDeviceToHost, st));
        records_to_host(x, a, nullptr, 1, B->nlq_rec);  // is_contain_N
    }
    CURC_HIP(stream_wait(st));
    return out;

}

// The helper context of c (settings copied from c), or null if it cannot be made.
curc_ctx* helper_ctx(curc_ctx* c) {
    if (!c->helper) {
        c->helper = curc_ctx_create(c->device);
        if (!c->helper) return nullptr;
        c->helper->side_copies = false;
    }
    curc_ctx* h = c->helper;
    h->timing = c->timing;
    h->index_csr = c->index_csr;
    h->tune = c->tune;
    h->graphs.enabled = c->graphs.enabled;
    return h;
}

// Malformed input (reported as CURC_EINVAL).
struct InputError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// FASTQ bytes of one block in device memory: file 1, and file 2 of a
// paired block (reads interleave as 2k / 2k+1).  When `ready` is set the
// stream waits on it before the first read of the bytes; `done`, when set,
// is recorded after the last one.
struct FqInput {
    const uint8_t* fq[2] = {nullptr, nullptr};
    uint64_t bytes[2] = {0, 0};
    // records parsed on the host: their sequences, n_ascii rows of L bytes
    // in read id order, in device memory (fq then unused)
    const uint8_t* ascii = nullptr;
    uint64_t n_ascii = 0;
    hipEvent_t ready = nullptr;
    hipEvent_t done = nullptr;
};

// Record starts of a block's FASTQ bytes (preprocess.cpp:147-193): returns
// the read count and leaves the start of each read's sequence line (byte
// offset in its file) in *seq.  Scratch: tile counts in `tile_slot`, the
// starts in `seq_slot`.  Throws InputError on malformed input.
uint64_t fq_parse(curc_ctx* c, const FqInput& in, uint64_t** seq, int tile_slot, int seq_slot, uint32_t* flags,
                  uint64_t* tile_buf = nullptr, uint64_t tile_cap = 0, uint64_t* seq_buf = nullptr,
                  uint64_t seq_cap = 0) {
    Scratch& S = c->S;
    hipStream_t st = c->st;
    const int nf = in.fq[1] ? 2 : 1;
    uint64_t tiles[2] = {0, 0}, off[2] = {0, 0};
    uint64_t total_tiles = 0;
    for (int f = 0; f < nf; ++f) {
        if (((uintptr_t)in.fq[f] & 15) != 0) throw InputError("FASTQ bytes must be 16-byte aligned on the device");
        tiles[f] = std::max<uint64_t>(1, (in.bytes[f] + FQ_TILE - 1) / FQ_TILE);
        off[f] = total_tiles;
        total_tiles += tiles[f] + 1;  // + the file's total
    }
    if (tile_buf && total_tiles > tile_cap) throw std::length_error("FASTQ block larger than the scratch buffer");
    uint64_t* tc = tile_buf ? tile_buf : S.as<uint64_t>(tile_slot, total_tiles);
    set_cnt(c, flags, 0u);
    for (int f = 0; f < nf; ++f) {
        if (in.bytes[f]) hipLaunchKernelGGL(k_fq_ends, dim3(1), dim3(1), 0, st, in.fq[f], in.bytes[f], flags);
        hipLaunchKernelGGL(k_fq_count, dim3((uint32_t)tiles[f]), dim3(256), 0, st, in.fq[f], in.bytes[f], tc + off[f]);
        hipLaunchKernelGGL(k_fq_scan, dim3(1), dim3(1024), 0, st, tc + off[f], tiles[f], tc + off[f] + tiles[f]);
    }
    uint64_t lines[2] = {0, 0};
    uint32_t fl = 0;
    for (int f = 0; f < nf; ++f)
        CURC_HIP(hipMemcpyAsync(&lines[f], tc + off[f] + tiles[f], 8, hipMemcpyDeviceToHost, st));
    CURC_HIP(hipMemcpyAsync(&fl, flags, 4, hipMemcpyDeviceToHost, st));
    CURC_HIP(stream_wait(st));
    if (fl & 1u) throw InputError("FASTQ block does not start with '@'");
    for (int f = 0; f < nf; ++f)
        if ((fl & 2u) || lines[f] % 4) throw InputError("truncated FASTQ record");
    const uint64_t n1 = lines[0] / 4;
    if (nf == 2 && lines[1] / 4 != n1) throw InputError("paired-end files hold different record counts");
    const uint64_t N = n1 * nf;
    if (N >= (uint64_t)INV) throw std::length_error("a block holds at most 2^32-2 reads");
    if (seq_buf && N > seq_cap) throw std::length_error("more FASTQ records than the caller's arrays hold");
    uint64_t* sq = seq_buf ? seq_buf : S.as<uint64_t>(seq_slot, N + 1);
    for (int f = 0; f < nf; ++f)
        if (in.bytes[f])
            hipLaunchKernelGGL(k_fq_lines, dim3((uint32_t)tiles[f]), dim3(256), 0, st, in.fq[f], in.bytes[f],
                               (const uint64_t*)(tc + off[f]), sq, (uint32_t)nf, (uint32_t)f);
    *seq = sq;
    return N;
}

// Parse + pack of one file's FASTQ bytes on the device into the caller's
// arrays (packed: max_reads * W words, isN optional, seq: max_reads starts);
// counters in cnt (device, C_COUNT words); tile counts in tile_buf when
// given (capacity tile_cap words), else in scratch.  Returns the read count.
uint64_t fq_pack(curc_ctx* c, const FqInput& in, uint32_t L, uint32_t W, uint64_t max_reads, uint64_t* packed,
                 uint8_t* isN, uint64_t* seq, uint32_t* cnt, uint64_t* tile_buf, uint64_t tile_cap) {
    hipStream_t st = c->st;
    fill<uint32_t>(c, cnt, C_COUNT, 0u);
    if (in.bytes[0] == 0) return 0;
    uint64_t* sq = nullptr;
    const uint64_t N = fq_parse(c, in, &sq, S_FQTILE, S_FQSEQ, cnt + C_FQERR, tile_buf, tile_cap, seq, max_reads);
    if (N) {
        if (isN) fill<uint8_t>(c, isN, N, (uint8_t)0);
        const FqReads fr{in.fq[0], in.fq[0], in.bytes[0], 0, sq, 0u};
        hipLaunchKernelGGL(k_pack_fq, dim3(grid_for(N * W, 1u << 20)), dim3(256), 0, st, fr, N, L, W, packed, isN,
                           cnt + C_FQERR);
    }
    uint32_t e = 0;
    CURC_HIP(hipMemcpyAsync(&e, cnt + C_FQERR, 4, hipMemcpyDeviceToHost, st));
    CURC_HIP(stream_wait(st));
    if (e) throw InputError("read length isn't fix");  // preprocess.cpp:154
    return N;
}

void run_block(curc_ctx* c, const char* ascii, uint64_t N, uint32_t L, int stop, uint32_t flags,
               curc_block* B, const FqInput* fq = nullptr) {
    curc_params P;
    curc_make_params(L, &P);
    const uint32_t W = P.W;
    Scratch& S = c->S;
    hipStream_t st = c->st;
    const bool keep = (flags & CURC_RUN_KEEP_STAGES) != 0;
    B->L = L;
    B->W = W;
    uint32_t* cnt = S.as<uint32_t>(S_CNT, C_COUNT);
    AsmOut hq_asm, lq_asm, nlq_asm;
    fill<uint32_t>(c, cnt, C_COUNT, 0u);

    // ---- a1/a2: record parse, N split + pack (preprocess.cpp:147-199, compress.cu:2151-2180)
    const uint8_t* d_ascii = nullptr;
    FqReads fr{};
    // FASTQ bytes parsed here on the device (else sequences parsed on the host)
    const bool fqd = fq && !fq->ascii;
    if (fq && fq->ascii) {
        if (fq->ready) CURC_HIP(hipStreamWaitEvent(st, fq->ready, 0));
        d_ascii = fq->ascii;
        N = fq->n_ascii;
    } else if (fq) {
        if (fq->ready) CURC_HIP(hipStreamWaitEvent(st, fq->ready, 0));
        Region r(c, "parse");
        uint64_t* seq = nullptr;
        N = fq_parse(c, *fq, &seq, S_FQTILE, S_FQSEQ, cnt + C_FQERR);
        fr = FqReads{fq->fq[0], fq->fq[1] ? fq->fq[1] : fq->fq[0], fq->bytes[0], fq->bytes[1], seq,
                     fq->fq[1] ? 1u : 0u};
        c->add_bytes("parse", 2.0 * (fq->bytes[0] + fq->bytes[1]) + 8.0 * N, (double)N);
    } else if (flags & CURC_RUN_INPUT_ON_DEVICE) {
        d_ascii = (const uint8_t*)ascii;
    } else {
        uint8_t* a = S.as<uint8_t>(S_ASCII, N * L + 16);
        h2d(c, a, (const uint8_t*)ascii, N * L);
        d_ascii = a;
    }
    B->n_reads = N;
    if (fq && N == 0) {
        // no records: an empty block
        if (fq->done) CURC_HIP(hipEventRecord(fq->done, st));
        return;
    }
    uint64_t* packed = S.as<uint64_t>(S_PACKED, N * W + 1);
    uint8_t* isN = S.as<uint8_t>(S_ISN, N + 1);
    uint32_t* nonN = S.as<uint32_t>(S_NONN, N + 1);
    uint32_t* Nids = S.as<uint32_t>(S_NIDS, N + 1);
    uint32_t n, nN;
    {
        Region r(c, "pack");
        fill<uint64_t>(c, packed + N * W, 1, 0ull);
        // one thread per packed word; isN is set only for reads with an 'N'
        fill<uint8_t>(c, isN, N + 1, (uint8_t)0);
        if (N && fqd)
            hipLaunchKernelGGL(k_pack_fq, dim3(grid_for(N * W, 1u << 20)), dim3(256), 0, st, fr, N, L, W, packed,
                               isN, cnt + C_FQERR);
        else if (N)
            hipLaunchKernelGGL(k_pack_w, dim3(grid_for(N * W, 1u << 20)), dim3(256), 0, st, d_ascii, N, L, W, packed,
                               isN);
        set_cnt(c, cnt + C_N, (uint32_t)N);
        compact(cnt + C_N, PredIsN{isN, false}, EmitIndex{nonN}, cnt + C_NONN, S, S_PRIM, st, N);
        compact(cnt + C_N, PredIsN{isN, true}, EmitIndex{Nids}, cnt + C_NN, S, S_PRIM, st, N);
        c->add_bytes("pack", (double)N * (L + 8.0 * W) + 4.0 * N, (double)N);
    }
    {
        uint32_t h[C_COUNT];
        CURC_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, st));
        CURC_HIP(stream_wait(st));
        if (fqd && h[C_FQERR]) throw InputError("read length isn't fix");  // preprocess.cpp:154
        n = h[C_NONN];
        nN = h[C_NN];
    }
    uint8_t* Nascii = S.as<uint8_t>(S_NASCII, (uint64_t)nN * L + 16);
    if (nN && fqd)
        hipLaunchKernelGGL(k_gather_ascii_fq, dim3(grid_for((uint64_t)nN * L)), dim3(256), 0, st, fr,
                           (const uint32_t*)Nids, (uint64_t)nN, L, Nascii);
    else if (nN)
        hipLaunchKernelGGL(k_gather_ascii, dim3(grid_for((uint64_t)nN * L)), dim3(256), 0, st, d_ascii,
                           (const uint32_t*)Nids, (uint64_t)nN, L, Nascii);
    if (fq && fq->done) CURC_HIP(hipEventRecord(fq->done, st));

    // ---- a4: dedup (compress.cu:1396-1420)
    uint32_t* next = S.as<uint32_t>(S_NEXT, N + 1);
    uint32_t* prev = S.as<uint32_t>(S_PREV, N + 1);
    uint16_t* off = S.as<uint16_t>(S_OFF, N + 1);
    uint32_t* sorted;
    uint8_t* eqf = S.as<uint8_t>(S_DEQ, (uint64_t)n + 1);  // see k_dedup_link
    {
        Region r(c, "dedup");
        fill<uint32_t>(c, next, N, INV);
        fill<uint32_t>(c, prev, N, INV);
        fill<uint16_t>(c, off, N, INVOFF);
        uint64_t* k0 = S.as<uint64_t>(S_K0, (uint64_t)n + 1);
        uint64_t* k1 = S.as<uint64_t>(S_K1, (uint64_t)n + 1);
        uint32_t* v0 = S.as<uint32_t>(S_V0, (uint64_t)n + 1);
        uint32_t* v1 = S.as<uint32_t>(S_V1, (uint64_t)n + 1);
        CURC_HIP(hipMemcpyAsync(v0, nonN, (uint64_t)n * 4, hipMemcpyDeviceToDevice, st));
        // word w holds bases [32w, 32w + 32) (fewer in the last word); only
        // its top 2 * nbases bits carry data, in 8-bit digits from `begin`
        auto begin_bit = [&](uint32_t w) {
            const uint32_t nbases = (w == W - 1) ? (L - 32 * (W - 1)) : 32;
            return ((64 - 2 * nbases) / 8) * 8;
        };
        // every read by its first 16 bases (the top 32 bits of its first
        // word; stable, so ties keep ascending id): half the passes of the
        // whole word, and reads that share their first 16 bases are nearly
        // all reads of one locus, which share the next 16 as well
        const uint32_t b0 = begin_bit(0);
        const bool split = b0 < 32;  // bits [b0, 32) of the first word left for the runs
        const uint64_t kmask = split ? 0xFFFFFFFF00000000ull : ~0ull;
        hipLaunchKernelGGL(k_gather_word, dim3(grid_for(n)), dim3(256), 0, st, (const uint64_t*)packed,
                           (const uint32_t*)v0, (uint64_t)n, W, 0u, k0);
        radix_sort_pairs<uint64_t>(k0, v0, k1, v1, n, split ? 32u : b0, 64, S, S_PRIM, st);
        if ((W > 1 || split) && n > 1) {
            // reads in runs of equal sorted bits, by the rest of their words
            uint32_t* tpos = S.as<uint32_t>(S_TPOS, (uint64_t)n + 1);
            set_cnt(c, cnt + C_TMP, n);
            compact(cnt + C_TMP, PredTied{k0, n, kmask}, EmitIota{tpos}, cnt + C_TIE, S, S_PRIM, st, n);
            const uint32_t m = read_cnt(c, cnt + C_TIE);
            if (m > 0) {
                uint32_t* sid = S.as<uint32_t>(S_TSID, (uint64_t)m + 1);
                uint32_t* p0 = S.as<uint32_t>(S_TP0, (uint64_t)m + 1);
                uint32_t* p1 = S.as<uint32_t>(S_TP1, (uint64_t)m + 1);
                uint32_t* head = S.as<uint32_t>(S_THEAD, (uint64_t)m + 1);
                uint32_t* run = S.as<uint32_t>(S_TRUN, (uint64_t)m + 1);
                uint64_t* tk0 = S.as<uint64_t>(S_TK0, (uint64_t)m + 1);
                uint64_t* tk1 = S.as<uint64_t>(S_TK1, (uint64_t)m + 1);
                const dim3 gm(grid_for(m));
                hipLaunchKernelGGL(k_tie_setup, gm, dim3(256), 0, st, (const uint32_t*)tpos,
                                   (const uint32_t*)(cnt + C_TIE), (const uint64_t*)k0, kmask, (const uint32_t*)v0,
                                   sid, p0, head);
                scan_excl(head, run, m, S, S_PRIM, st);
                hipLaunchKernelGGL(k_tie_runid, gm, dim3(256), 0, st, (const uint32_t*)(cnt + C_TIE),
                                   (const uint32_t*)head, (const uint32_t*)run, run);
                uint32_t* perm = p0;
                uint32_t* permt = p1;
                for (uint32_t w = W - 1; w >= 1; --w) {
                    hipLaunchKernelGGL(k_gather_word_perm, gm, dim3(256), 0, st, (const uint64_t*)packed,
                                       (const uint32_t*)sid, (const uint32_t*)perm, (uint64_t)m, W, w, tk0);
                    radix_sort_pairs<uint64_t>(tk0, perm, tk1, permt, m, begin_bit(w), 64, S, S_PRIM, st);
                }
                if (split) {
                    // the first word's remaining bits, most significant of the words after it
                    hipLaunchKernelGGL(k_gather_word_perm, gm, dim3(256), 0, st, (const uint64_t*)packed,
                                       (const uint32_t*)sid, (const uint32_t*)perm, (uint64_t)m, W, 0u, tk0);
                    radix_sort_pairs<uint64_t>(tk0, perm, tk1, permt, m, b0, 32, S, S_PRIM, st);
                }
                // then by run number (stable): runs back in position order
                uint32_t* rk0 = reinterpret_cast<uint32_t*>(tk0);
                uint32_t* rk1 = reinterpret_cast<uint32_t*>(tk1);
                hipLaunchKernelGGL(k_gather_u32_perm, gm, dim3(256), 0, st, (const uint32_t*)run,
                                   (const uint32_t*)perm, (uint64_t)m, rk0);
                radix_sort_pairs<uint32_t>(rk0, perm, rk1, permt, m, 0, std::max(8u, 8 * ((log2u(m) + 7) / 8)),
                                           S, S_PRIM, st);
                hipLaunchKernelGGL(k_tie_scatter, gm, dim3(256), 0, st, (const uint32_t*)tpos, (const uint32_t*)sid,
                                   (const uint32_t*)perm, (uint64_t)m, v0);
            }
        }
        sorted = v0;
        if (n > 1)
            hipLaunchKernelGGL(k_dedup_link, dim3(grid_for(n)), dim3(256), 0, st, (const uint64_t*)packed,
                               (const uint32_t*)sorted, (uint64_t)n, W, next, prev, off, eqf);
        else
            fill<uint8_t>(c, eqf, 1, (uint8_t)0);
        c->add_bytes("dedup", (double)n * 8 * W + 2.0 * n * (8 * W + 4), (double)n);
    }
    // keep the sorted ids in a stable slot (v0/v1 are reused by later sorts)
    uint32_t* srt = S.as<uint32_t>(S_SORTED, (uint64_t)n + 1);
    CURC_HIP(hipMemcpyAsync(srt, sorted, (uint64_t)n * 4, hipMemcpyDeviceToDevice, st));
    sorted = srt;
    set_cnt(c, cnt + C_TMP, n);
    if (keep) {
        d2h(c, B->next_dedup, next, N);
        d2h(c, B->prev_dedup, prev, N);
    }
    if (keep) {
        d2h(c, B->sorted, sorted, n);
        d2h(c, B->nonN, nonN, n);
    }
    d2h(c, B->Nids, Nids, nN);
    if (stop <= CURC_STAGE_DEDUP) {
        d2h(c, B->next, next, N);
        d2h(c, B->prev, prev, N);
        d2h(c, B->offset, off, N);
        CURC_HIP(stream_wait(st));
        return;
    }

    bool ov_counted = false;
    // ---- a5: prefix k-mer index (compress.cu:1427-1471)
    uint32_t* va = S.as<uint32_t>(S_VA, (uint64_t)n + 1);
    uint32_t* sl0 = S.as<uint32_t>(S_SL0, (uint64_t)n + 1);
    // the prefix list (reads without a predecessor) and the suffix list
    // (without a successor), in dedup order: one single-pass launch for both
    // when they fit its tile limit
    uint32_t lb_err = 0;
    compact_pair(c, cnt + C_TMP, n, PredNoPred{eqf}, EmitGather{sorted, va}, cnt + C_NPRE, PredNoSucc{eqf},
                 EmitGather{sorted, sl0}, cnt + C_NSUF, &lb_err);
    const uint32_t npre = read_cnt(c, cnt + C_NPRE);
    const uint32_t nsuf = read_cnt(c, cnt + C_NSUF);
    if (lb_err != 0) throw std::runtime_error("prefix / suffix list compaction: look-back did not complete");
    {
        const uint32_t H = hash_size(npre);
        const uint32_t mask = H - 1;
        uint32_t* kr = S.as<uint32_t>(S_KR, (uint64_t)H + 2);
        uint64_t* vk = S.as<uint64_t>(S_VK, (uint64_t)npre + 1);
        // bucket slots up to 2^28 buckets (16 GiB), the CSR arrays alone beyond
        // and for sparse tables (fewer than H / 8 prefixes)
        PrefixSlot* pslots = (H <= PSLOT_MAX_H && !c->index_csr && (uint64_t)npre * 8 >= H)
                                 ? S.as<PrefixSlot>(S_PSLOT, (uint64_t)H)
                                 : nullptr;
        // prefix k-mer filter: >= filter_bits bits per prefix (2^20..2^32
        // bits), so a probe without an equal k-mer usually stops at one bit.
        // With verdicts (overlap probe form 2: reads of at most 256 bases,
        // k >= OVM + 4) the filter is minimizer-blocked and consulted once
        // per suffix for all offsets; otherwise it is the plain bitmap.
        uint64_t fbits = 1ull << 20;
        while (fbits < (uint64_t)c->tune.filter_bits * npre && fbits < (1ull << 32)) fbits <<= 1;
        const uint32_t fmask = (uint32_t)(fbits - 1);
        const uint64_t fwords = fbits / 32;
        uint32_t* pfilt = S.as<uint32_t>(S_PFILT, fwords);
        const bool verd = c->tune.ov_queue >= 2 && W <= 8 && P.k >= OVM + 4;
        const uint32_t bmask = (uint32_t)(fbits / (32 * OVB_WORDS) - 1);
        {
            Region r(c, "prefix_index");
            uint32_t* hk0 = S.as<uint32_t>(S_MK0, (uint64_t)npre + 1);
            uint32_t* hk1 = S.as<uint32_t>(S_MK1, (uint64_t)npre + 1);
            uint32_t* vb = S.as<uint32_t>(S_MV1, (uint64_t)npre + 1);
            uint32_t* va_ = va;
            fill<uint32_t>(c, pfilt, fwords, 0u);
            hipLaunchKernelGGL(k_prefix_hash, dim3(grid_for(npre)), dim3(256), 0, st, (const uint64_t*)packed,
                               (const uint32_t*)va, (uint64_t)npre, W, P.kip, P.k, mask, hk0,
                               verd ? nullptr : pfilt, fmask, verd ? pfilt : nullptr, bmask);
            radix_sort_pairs<uint32_t>(hk0, va_, hk1, vb, npre, 0, log2u(H), S, S_PRIM, st);
            if (va_ != va) CURC_HIP(hipMemcpyAsync(va, va_, (uint64_t)npre * 4, hipMemcpyDeviceToDevice, st));
            hipLaunchKernelGGL(k_gap_fill, dim3(grid_for((uint64_t)npre + 1)), dim3(256), 0, st,
                               (const uint32_t*)hk0, (uint64_t)npre, (uint64_t)H + 1, kr);
            hipLaunchKernelGGL(k_gather_key0, dim3(grid_for(npre)), dim3(256), 0, st, (const uint64_t*)packed,
                               (const uint32_t*)va, (uint64_t)npre, W, vk);
            if (pslots)
                hipLaunchKernelGGL(k_prefix_slots, dim3(grid_for(H)), dim3(256), 0, st, (const uint32_t*)kr,
                                   (const uint32_t*)va, (const uint64_t*)vk, (uint64_t)H, pslots);
            c->add_bytes("prefix_index", (double)npre * (16 + 8 + 16) + 8.0 * H, (double)npre);
        }
        // ---- a6: greedy overlap loop (compress.cu:1473-1532)
        {
            Region r(c, "overlap");
            uint32_t* sl[1] = {sl0};
            uint32_t* choice = S.as<uint32_t>(S_CHOICE, (uint64_t)nsuf + 1);
            uint32_t* grp = S.as<uint32_t>(S_GRP, (uint64_t)nsuf + 1);
            uint32_t* memb = S.as<uint32_t>(S_MEMB, (uint64_t)nsuf + 1);
            uint32_t* resl = S.as<uint32_t>(S_RESL, (uint64_t)nsuf + 1);
            uint32_t* pool = S.as<uint32_t>(S_POOL, (uint64_t)nsuf + 1);
            uint32_t* owner = S.as<uint32_t>(S_OWNER, (uint64_t)npre + 1);
            uint32_t* cont = S.as<uint32_t>(S_CONT, (uint64_t)npre + 1);
            uint32_t* head = S.as<uint32_t>(S_HEAD, (uint64_t)npre + 1);
            fill<uint32_t>(c, owner, (uint64_t)npre + 1, INV);
            fill<uint32_t>(c, head, (uint64_t)npre + 1, INV);
            fill<uint32_t>(c, cont, (uint64_t)npre + 1, 0u);
            // one static suffix list for all offsets (its length is nsuf, the
            // host value of cnt[C_NSUF]); linked suffixes are skipped by flag
            const uint32_t g = grid_for(nsuf, c->tune.ov_grid);
            uint8_t* slinked = S.as<uint8_t>(S_SLINK, (uint64_t)nsuf + 1);
            fill<uint8_t>(c, slinked, (uint64_t)nsuf + 1, (uint8_t)0);
            // word-major copy of the suffixes' words for the probe forms that
            // read a word of every suffix at every offset (0 and 1)
            uint64_t* sw = nullptr;
            if (!verd) {
                sw = S.as<uint64_t>(S_SUFW, (uint64_t)nsuf * W + 1);
                if (nsuf)
                    hipLaunchKernelGGL(k_gather_soa, dim3(grid_for((uint64_t)nsuf * W)), dim3(256), 0, st,
                                       (const uint64_t*)packed, (const uint32_t*)sl[0], (uint64_t)nsuf, W, sw);
            }
            // filter verdicts of every offset, one 64-suffix word per offset
            const uint64_t vnw = ((uint64_t)nsuf + 63) / 64;
            uint64_t* vb = nullptr;
            if (verd && nsuf && P.max_off > 1) {
                vb = S.as<uint64_t>(S_OVVB, (uint64_t)(P.max_off - 1) * vnw);
                const dim3 vg(grid_for(nsuf, 65536));
#define CURC_OV_VERDICTS(WM)                                                                           \
    do {                                                                                               \
        if (P.k == 32)                                                                                 \
            hipLaunchKernelGGL((k_ov_verdicts<WM, true>), vg, dim3(256), 0, st, (const uint64_t*)packed,      \
                               (const uint32_t*)sl[0], nsuf, W, P.k, P.kip, P.max_off, (const uint32_t*)pfilt, \
                               bmask, vb, vnw, (const uint32_t*)nullptr);                              \
        else                                                                                           \
            hipLaunchKernelGGL((k_ov_verdicts<WM, false>), vg, dim3(256), 0, st, (const uint64_t*)packed,     \
                               (const uint32_t*)sl[0], nsuf, W, P.k, P.kip, P.max_off, (const uint32_t*)pfilt, \
                               bmask, vb, vnw, (const uint32_t*)nullptr);                              \
    } while (0)
                if (W <= 3)
                    CURC_OV_VERDICTS(3);
                else if (W <= 4)
                    CURC_OV_VERDICTS(4);
                else if (W <= 5)
                    CURC_OV_VERDICTS(5);
                else
                    CURC_OV_VERDICTS(8);
#undef CURC_OV_VERDICTS
            }
            // probe by verdict words (tuning ov_queue 3): workgroup b owns
            // verdict words [b*wper, b*wper + wper); linked suffixes are
            // tracked in the bitmap slbits
            const bool wscan = c->tune.ov_queue == 3 && vb;
            uint64_t* slbits = nullptr;
            uint32_t wper = 0, gw = 0;
            if (wscan) {
                slbits = S.as<uint64_t>(S_OVLIST, vnw + 1);
                fill<uint64_t>(c, slbits, vnw + 1, 0ull);
                const uint64_t want = (vnw + c->tune.ov_grid - 1) / c->tune.ov_grid;
                wper = (uint32_t)std::max<uint64_t>(OVW_TILE, (want + OVW_TILE - 1) / OVW_TILE * OVW_TILE);
                gw = (uint32_t)std::max<uint64_t>(1, (vnw + wper - 1) / wper);
            }
            // work counters of the loop (probes, bucket searches)
            uint64_t* ovw = S.as<uint64_t>(S_WORK, 4);
            fill<uint64_t>(c, ovw, 2, 0ull);
            // queued probe: workgroup b owns suffixes [b*per, b*per + per),
            // per a whole number of filter tiles, and lists its proposers in
            // that segment of qcand
            const bool queued = c->tune.ov_queue != 0 && !wscan;
            uint32_t per = 0, gq = 0;
            uint32_t *qcand = nullptr, *qn = nullptr;
            if (queued) {
                const uint64_t want = ((uint64_t)nsuf + c->tune.ov_grid - 1) / c->tune.ov_grid;
                per = (uint32_t)std::max<uint64_t>(OVQ_TILE, (want + OVQ_TILE - 1) / OVQ_TILE * OVQ_TILE);
                gq = (uint32_t)std::max<uint64_t>(1, ((uint64_t)nsuf + per - 1) / per);
                qcand = S.as<uint32_t>(S_OVCAND, (uint64_t)nsuf + 1);
                qn = S.as<uint32_t>(S_OVNC, (uint64_t)gq);
            }
            uint32_t* passed = queued ? S.as<uint32_t>(S_OVPASS, (uint64_t)gq) : nullptr;
            if (passed) fill<uint32_t>(c, passed, gq, 0u);
            if (wscan) {
                qcand = S.as<uint32_t>(S_OVCAND, (uint64_t)gw * wper * 64 + 1);
                qn = S.as<uint32_t>(S_OVNC, (uint64_t)gw);
                passed = S.as<uint32_t>(S_OVPASS, (uint64_t)gw);
                fill<uint32_t>(c, passed, gw, 0u);
            }
            auto loop = [&]() {
                for (uint32_t o = 1; o < P.max_off; ++o) {
                    ClaimState cs{packed, W, L, sl[0], cnt + C_NSUF, va, vk, cnt + C_NPRE, next, prev, off,
                                  choice, grp, owner, cont, head, memb, resl, cnt + C_RESN, pool, cnt + C_POOL,
                                  nullptr, slinked, nullptr};
                    if (wscan) {
                        cs.slbits = slbits;
#define CURC_OV_PROBE_W(WM)                                                                                    \
    hipLaunchKernelGGL((k_ov_probe_w<WM>), dim3(gw), dim3(256), 0, st, cs, (const uint32_t*)kr,                  \
                       (const PrefixSlot*)pslots, P.k, P.kip, mask, o, wper, (const uint64_t*)vb, vnw, qcand, qn,   \
                       passed)
                        if (W <= 3)
                            CURC_OV_PROBE_W(3);
                        else if (W <= 4)
                            CURC_OV_PROBE_W(4);
                        else if (W <= 5)
                            CURC_OV_PROBE_W(5);
                        else
                            CURC_OV_PROBE_W(8);
#undef CURC_OV_PROBE_W
                        hipLaunchKernelGGL(k_ov_finalize_q, dim3(gw), dim3(256), 0, st, cs, o, wper * 64,
                                           (const uint32_t*)qcand, (const uint32_t*)qn);
                    } else if (queued && vb) {
#define CURC_OV_PROBE_V(WM)                                                                                    \
    hipLaunchKernelGGL((k_ov_probe_q<true, WM>), dim3(gq), dim3(256), 0, st, cs, (const uint32_t*)kr,            \
                       (const PrefixSlot*)pslots, (const uint32_t*)nullptr, fmask, mask, (const uint64_t*)sw, P.k, \
                       P.kip, o, per, qcand, qn, (const uint64_t*)vb, vnw, passed)
                        if (W <= 3)
                            CURC_OV_PROBE_V(3);
                        else if (W <= 4)
                            CURC_OV_PROBE_V(4);
                        else if (W <= 5)
                            CURC_OV_PROBE_V(5);
                        else
                            CURC_OV_PROBE_V(8);
#undef CURC_OV_PROBE_V
                        hipLaunchKernelGGL(k_ov_finalize_q, dim3(gq), dim3(256), 0, st, cs, o, per,
                                           (const uint32_t*)qcand, (const uint32_t*)qn);
                    } else if (queued) {
                        hipLaunchKernelGGL((k_ov_probe_q<false, 1>), dim3(gq), dim3(256), 0, st, cs, (const uint32_t*)kr,
                                           (const PrefixSlot*)pslots, (const uint32_t*)pfilt, fmask, mask,
                                           (const uint64_t*)sw, P.k, P.kip, o, per, qcand, qn,
                                           (const uint64_t*)nullptr, (uint64_t)0, passed);
                        hipLaunchKernelGGL(k_ov_finalize_q, dim3(gq), dim3(256), 0, st, cs, o, per,
                                           (const uint32_t*)qcand, (const uint32_t*)qn);
                    } else {
                        hipLaunchKernelGGL(k_ov_probe, dim3(g), dim3(256), 0, st, cs, (const uint32_t*)kr,
                                           (const PrefixSlot*)pslots, (const uint32_t*)pfilt, fmask, mask,
                                           (const uint64_t*)sw, P.k, P.kip, o);
                        hipLaunchKernelGGL(k_claim_finalize, dim3(g), dim3(256), 0, st, cs, o);
                    }
                    hipLaunchKernelGGL((k_claim_resolve<false, Enc2>), dim3(4 * c->cus), dim3(256), 0, st, cs, o);
                }
            };
            GraphKey key;
            key.v[0] = 2;  // the overlap loop
            key.v[1] = W;
            key.v[2] = L;
            key.v[3] = nsuf;
            key.v[4] = mask;
            key.v[5] = fmask;
            key.v[6] = pslots != nullptr;
            key.v[7] = (uint64_t)packed;
            key.v[8] = (uint64_t)next;
            key.v[9] = (uint64_t)prev;
            key.v[10] = (uint64_t)off;
            key.v[11] = (uint64_t)cnt;
            key.v[12] = ((uint64_t)queued << 32) | (queued ? per : g);
            key.v[13] = (uint64_t)vb | (wscan ? 1u : 0u);
            key.v[14] = (uint64_t)slbits;
            key.v[15] = ((uint64_t)gw << 32) | wper;
            key.v[16] = (uint64_t)qcand;
            key.v[17] = (uint64_t)qn;
            key.v[18] = (uint64_t)passed;
            c->graphs.run(key, S, st, loop);
            // work of the loop: probes (suffix x offset) and bucket searches
            // (probes that passed the k-mer filter); read back with the
            // partition's counts
            if (nsuf)
                hipLaunchKernelGGL(k_ov_probe_count, dim3(grid_for(nsuf, 1024)), dim3(256), 0, st, (const uint32_t*)sl[0],
                                   nsuf, (const uint16_t*)off, P.max_off, ovw);
            if (passed)
                hipLaunchKernelGGL(k_sum_u32, dim3(grid_for(wscan ? gw : gq)), dim3(256), 0, st,
                                   (const uint32_t*)passed, wscan ? gw : gq, ovw + 1);
            CURC_HIP(hipMemcpyAsync(c->ov_work, ovw, 16, hipMemcpyDeviceToHost, st));
            ov_counted = true;
        }
    }
    if (keep) {
        d2h(c, B->next_ov, next, N);
        d2h(c, B->prev_ov, prev, N);
        d2h(c, B->off_ov, off, N);
    }
    if (stop <= CURC_STAGE_OVERLAP) {
        d2h(c, B->next, next, N);
        d2h(c, B->prev, prev, N);
        d2h(c, B->offset, off, N);
        CURC_HIP(stream_wait(st));
        return;
    }

    // ---- a7: prune + stable partition (compress.cu:1539-1572)
    uint32_t* hq = S.as<uint32_t>(S_HQ, (uint64_t)n + 1);
    uint32_t* unm = S.as<uint32_t>(S_UNM, (uint64_t)n + 1);
    uint32_t n_hq, n_unm;
    {
        Region r(c, "prune");
        uint8_t* res = S.as<uint8_t>(S_RES, N + 1);
        hipLaunchKernelGGL(k_prune_mark, dim3(grid_for(N)), dim3(256), 0, st, (const uint32_t*)next,
                           (const uint32_t*)prev, (const uint16_t*)off, N, res);
        hipLaunchKernelGGL(k_prune_unlink, dim3(grid_for(N)), dim3(256), 0, st, next, prev, (const uint8_t*)res, N);
        set_cnt(c, cnt + C_TMP, n);
        // both lists of the partition in one single-pass launch
        compact_pair(c, cnt + C_TMP, n, PredRes{sorted, res, true}, EmitGather{sorted, hq}, cnt + C_HQ,
                     PredRes{sorted, res, false}, EmitGather{sorted, unm}, cnt + C_UNM, &lb_err);
        c->add_bytes("prune", 28.0 * N, (double)N);
    }
    n_hq = read_cnt(c, cnt + C_HQ);
    n_unm = read_cnt(c, cnt + C_UNM);
    if (lb_err != 0) throw std::runtime_error("HQ / unmatched partition: look-back did not complete");
    if (ov_counted)  // per probe: the suffix's k-mer, the bucket's bounds; per search: a prefix's id and words
        c->add_bytes("overlap", (double)c->ov_work[0] * 32 + (double)c->ov_work[1] * (4 + 8.0 * W),
                     (double)c->ov_work[0]);
    if (keep) {
        d2h(c, B->next_pr, next, N);
        d2h(c, B->prev_pr, prev, N);
    }
    if (keep || stop <= CURC_STAGE_PRUNE) {
        d2h(c, B->hq, hq, n_hq);
        d2h(c, B->unmatched, unm, n_unm);
    }
    if (stop <= CURC_STAGE_PRUNE) {
        d2h(c, B->next, next, N);
        d2h(c, B->prev, prev, N);
        d2h(c, B->offset, off, N);
        CURC_HIP(stream_wait(st));
        return;
    }

    // ---- a8: exact all-offset matching of HQ chain ends (compress.cu:1574-1591)
    {
        Region r(c, "full_match");
        uint32_t* pre = S.as<uint32_t>(S_PRE, (uint64_t)n_hq + 1);
        uint32_t* suf = S.as<uint32_t>(S_SUF, (uint64_t)n_hq + 1);
        set_cnt(c, cnt + C_TMP, n_hq);
        uint32_t lb_err = 0;
        compact_pair(c, cnt + C_TMP, n_hq, PredLinkState{hq, prev}, EmitGather{hq, pre}, cnt + C_PRE,
                     PredLinkState{hq, next}, EmitGather{hq, suf}, cnt + C_SUF, &lb_err);
        CURC_HIP(stream_wait(st));
        if (lb_err != 0) throw std::runtime_error("HQ list compaction: look-back did not complete");
        full_match<Enc2>(c, packed, W, L, pre, suf, cnt, n_hq, next, prev, off, "full_match");
    }
    if (keep || stop <= CURC_STAGE_FULLMATCH) {
        d2h(c, B->next, next, N);
        d2h(c, B->prev, prev, N);
        d2h(c, B->offset, off, N);
        CURC_HIP(stream_wait(st));
    }
    if (stop <= CURC_STAGE_FULLMATCH) return;

    // ---- a9: pseudogenome assembly (compress.cu:1613-1618)
    uint64_t* hq_pg;
    {
        Region r(c, "assembly");
        hq_asm = assemble(c, hq, n_hq, N, next, prev, off, L, cnt, S_PGIDS);
        const AsmOut& a = hq_asm;
        B->hq_len = a.len;
        hq_pg = build_pg(c, packed, W, a, S_PG_HQ);
        // records and text go to the host while mapping runs (own slots:
        // nothing later in the block writes them)
        curc_alignment* drec = S.as<curc_alignment>(S_RECHQ, (uint64_t)a.n + 1);
        if (a.n)
            hipLaunchKernelGGL(k_rec_aos, dim3(grid_for(a.n)), dim3(256), 0, st, (const uint32_t*)a.rid,
                               (const uint64_t*)a.pos, (const uint32_t*)nullptr, (uint64_t)a.n, 2u, 0ull, drec);
        d2h_side(c, B->hq_rec, (const curc_alignment*)drec, a.n, 0);
        d2h_side(c, B->hq_pg2, (const uint64_t*)hq_pg, (a.len + 31) / 32 + 1, 0);
    }
    if (stop <= CURC_STAGE_ASSEMBLY) return;

    // ---- a10: mapping (compress.cu:1620-1640, 268-595)
    {
        const bool learned = (flags & CURC_RUN_LEARNED) && c->has_scorer;
        const bool trace = (flags & CURC_RUN_TRACE) != 0;
        const bool phased = learned || trace;
        if (trace) {
            S.as<unsigned long long>(S_TRCNT, 1);
            fill<unsigned long long>(c, S.as<unsigned long long>(S_TRCNT, 1), 1, 0ull);
        }
        const uint64_t ref_len = B->hq_len;
        const uint64_t nm = (uint64_t)n_unm + nN;
        uint8_t* last = S.as<uint8_t>(S_LAST, nm + 1);
        uint64_t* word = S.as<uint64_t>(S_WORD, nm + 1);
        if (nm) hipLaunchKernelGGL(k_map_init, dim3(grid_for(nm)), dim3(256), 0, st, last, word, (uint64_t)n_unm, nm, P.max_mm);
        uint64_t index_count = 0;
        if (ref_len + 1 >= MAP_K) index_count = (ref_len - MAP_K + 1 + MAP_REF_STEP - 1) / MAP_REF_STEP;
        // the N-reads' masks of bases outside A, C, G, T: every mapping kernel
        // then compares N-reads word by word, like the other reads
        uint64_t* nmask = S.as<uint64_t>(S_NMASK, (uint64_t)nN * W + 1);
        if (nN)
            hipLaunchKernelGGL(k_nmask, dim3(grid_for((uint64_t)nN * W)), dim3(256), 0, st, (const uint8_t*)Nascii,
                               (uint64_t)nN, L, W, nmask);
        // work counters of a mapping pass, when timing (stats entries named
        // map_*), and its algorithmic bytes: each active read's words, its
        // state write (1 + 8 B), 16 B per k-mer probe (two CSR offsets and
        // the hash of a register-resident k-mer) and, per verified
        // candidate, its 4 B position plus the P bytes of pseudogenome
        auto map_stats = [&](const char* region, unsigned long long* work) {
            if (!c->timing) return;
            unsigned long long wk[8];
            CURC_HIP(hipMemcpyAsync(wk, work, sizeof(wk), hipMemcpyDeviceToHost, st));
            CURC_HIP(stream_wait(st));
            const double P8 = 8.0 * W;
            c->add_bytes(region, (double)wk[0] * (P8 + 9) + (double)wk[1] * 16 + (double)wk[2] * (4 + P8),
                         (double)wk[2]);
            const char* names[6] = {"map_active", "map_probes", "map_cands", "map_first_probe", "map_open_probes",
                                    "map_open_cands"};
            for (int q = 0; q < 6; ++q) c->add_bytes(names[q], 0.0, (double)wk[q]);
        };
        const char* phased_region = learned ? "mapping_learned" : "mapping_traced";
        // both strands' buckets in one index: bucket-inline slots, one index
        // part, reads of <= 256 bases.  The plain mode runs both strand passes
        // in one launch (k_map_both); the phased modes (compact slots only)
        // run their probe ranges strand by strand on it.
        const uint32_t H1 = index_count ? hash_size(index_count) : 0;
        const bool both = c->tune.map_both && (!phased || c->tune.map_both == 2) && L >= MAP_K && nm > 0 && W <= 8 &&
                          !c->index_csr &&
                          index_count > 0 && index_count <= (1ull << 30) && H1 <= MAP_SLOT_MAX_H &&
                          index_count * 8 >= H1 && ref_len < (1ull << 32);  // 32-bit positions
        if (both) {
            const uint64_t count = index_count;
            const uint32_t H = H1;
            // compact slots (tuning map_both 2) or 128-byte lines (1)
            const bool cs = c->tune.map_both == 2;
            uint32_t* bk = cs ? reinterpret_cast<uint32_t*>(S.as<uint4>(S_MBK, 2 * (uint64_t)H))
                              : S.as<uint32_t>(S_MBK, (uint64_t)H * MAP_SLOT2);
            uint32_t* ovf = nullptr;
            {
                Region r(c, "map_index");
                // each strand's (hash, position) pairs sorted, then both slots
                // of every bucket written together
                const int sk[2][5] = {{S_MK0, S_MK1, S_MV0, S_MV1, S_KRF}, {S_MK2, S_MK3, S_MV2, S_MV3, S_KRF2}};
                const uint32_t* sv[2];
                const uint32_t* sr[2];
                for (int strand = 0; strand < 2; ++strand) {
                    uint32_t* k0 = S.as<uint32_t>(sk[strand][0], count + 1);
                    uint32_t* k1 = S.as<uint32_t>(sk[strand][1], count + 1);
                    uint32_t* v0 = S.as<uint32_t>(sk[strand][2], count + 1);
                    uint32_t* v1 = S.as<uint32_t>(sk[strand][3], count + 1);
                    uint32_t* krf = S.as<uint32_t>(sk[strand][4], (uint64_t)H + 2);
                    hipLaunchKernelGGL(k_map_hash, dim3(grid_for(count)), dim3(256), 0, st, (const uint64_t*)hq_pg,
                                       (uint64_t)0, count, strand == 1, H - 1, k0, v0);
                    radix_sort_pairs<uint32_t>(k0, v0, k1, v1, count, 0, log2u(H), S, S_PRIM, st);
                    hipLaunchKernelGGL(k_gap_fill, dim3(grid_for(count + 1)), dim3(256), 0, st,
                                       (const uint32_t*)k0, count, (uint64_t)H + 1, krf);
                    sv[strand] = v0;
                    sr[strand] = krf;
                }
                if (cs) {
                    // overflow offsets by a scan of the overflow sizes (the CSR
                    // slot S_MKR is free on this path), overflow positions in
                    // S_MPOS, at most all positions of both strands
                    uint32_t* ovc = S.as<uint32_t>(S_MKR, (uint64_t)H + 2);
                    ovf = S.as<uint32_t>(S_MPOS, 2 * count + 2);
                    hipLaunchKernelGGL(k_cslot_count, dim3(grid_for((uint64_t)H + 1)), dim3(256), 0, st, sr[0],
                                       sr[1], (uint64_t)H, ovc);
                    scan_excl(ovc, ovc, (uint64_t)H + 1, S, S_PRIM, st);
                    hipLaunchKernelGGL(k_cslot_write, dim3(grid_for((uint64_t)H)), dim3(256), 0, st, sv[0], sr[0],
                                       sv[1], sr[1], (uint64_t)H, (const uint32_t*)ovc, reinterpret_cast<uint4*>(bk),
                                       ovf);
                } else {
                    hipLaunchKernelGGL(k_map_slots2, dim3(grid_for((uint64_t)H * 8, 1u << 20)), dim3(256), 0, st,
                                       sv[0], sr[0], sv[1], sr[1], (uint64_t)H, bk);
                }
                c->add_bytes("map_index", 2.0 * ((double)count * (16 + 8) + 8.0 * H), 2.0 * (double)count);
            }
            unsigned long long* work = S.as<unsigned long long>(S_MWORK, 8);
            MapArgs a{hq_pg, ref_len, nullptr, nullptr, H - 1, 0, packed, W, L, unm, n_unm,
                      Nascii, nN, last, word, work, P.max_mm, bk};
            a.bk_stride = MAP_SLOT2;
            a.bk_off = 0;
            a.cs = cs ? 1u : 0u;
            a.ovf = ovf;
            a.nids = Nids;
            a.nmask = nmask;
            if (phased) {
                for (int strand = 0; strand < 2; ++strand) {
                    fill<unsigned long long>(c, work, 8, 0ull);
                    {
                        Region r(c, phased_region);
                        map_phased(c, a, strand == 1, nm, cnt, learned, trace);
                    }
                    map_stats(phased_region, work);
                }
            } else {
                fill<unsigned long long>(c, work, 8, 0ull);
                {
                    Region r(c, "mapping");
                    // the N-reads' launch (a few percent of the reads, each
                    // probing long) on the helper context's idle stream beside
                    // the other reads' launch, instead of after it
                    curc_ctx* hx = (nN && c->side_copies) ? helper_ctx(c) : nullptr;
                    if (hx) {
                        for (auto& e : c->mev)
                            if (!e) CURC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                        CURC_HIP(hipEventRecord(c->mev[0], st));
                        CURC_HIP(hipStreamWaitEvent(hx->st, c->mev[0], 0));
                        launch_map_both(W, c->tune.map_grid, st, a, hx->st);
                        CURC_HIP(hipEventRecord(c->mev[1], hx->st));
                        CURC_HIP(hipStreamWaitEvent(st, c->mev[1], 0));
                    } else {
                        launch_map_both(W, c->tune.map_grid, st, a, st);
                    }
                }
                // as for the separate passes, over both strands
                map_stats("mapping", work);
            }
        } else if (L >= MAP_K && nm > 0) {
            for (int strand = 0; strand < 2; ++strand) {
                const bool rc = strand == 1;
                uint64_t parts[2][2];
                int np = 1;
                if (index_count > (1ull << 30)) {  // compress.cu:341-353
                    uint64_t c1 = index_count / 2;
                    parts[0][0] = 0;
                    parts[0][1] = c1;
                    parts[1][0] = c1 * MAP_REF_STEP;
                    parts[1][1] = index_count - c1;
                    np = 2;
                } else {
                    parts[0][0] = 0;
                    parts[0][1] = index_count;
                }
                for (int pi = 0; pi < np; ++pi) {
                    const uint64_t start = parts[pi][0], count = parts[pi][1];
                    const uint32_t H = hash_size(count);
                    // buckets inline in 64-byte slots up to 2^28 of them (16 GiB), CSR beyond
                    // and for sparse tables (fewer than H / 8 entries: the reference's
                    // minimum of 2^24 buckets would cost 1 GiB of slots for a small block)
                    const bool inline_bk = H <= MAP_SLOT_MAX_H && !c->index_csr && count * 8 >= H;
                    uint32_t* kr = inline_bk ? nullptr : S.as<uint32_t>(S_MKR, (uint64_t)H + 2);
                    uint32_t* pos = inline_bk ? nullptr : S.as<uint32_t>(S_MPOS, count + 2);
                    uint32_t* bk = inline_bk ? S.as<uint32_t>(S_MBK, (uint64_t)H * MAP_SLOT) : nullptr;
                    {
                        Region r(c, "map_index");
                        uint32_t* k0 = S.as<uint32_t>(S_MK0, count + 1);
                        uint32_t* k1 = S.as<uint32_t>(S_MK1, count + 1);
                        uint32_t* v0 = S.as<uint32_t>(S_MV0, count + 1);
                        uint32_t* v1 = S.as<uint32_t>(S_MV1, count + 1);
                        uint32_t* krf = S.as<uint32_t>(S_KRF, (uint64_t)H + 2);
                        if (count)
                            hipLaunchKernelGGL(k_map_hash, dim3(grid_for(count)), dim3(256), 0, st,
                                               (const uint64_t*)hq_pg, start, count, rc, H - 1, k0, v0);
                        radix_sort_pairs<uint32_t>(k0, v0, k1, v1, count, 0, log2u(H), S, S_PRIM, st);
                        hipLaunchKernelGGL(k_gap_fill, dim3(grid_for(count + 1)), dim3(256), 0, st,
                                           (const uint32_t*)k0, count, (uint64_t)H + 1, krf);
                        if (inline_bk) {
                            hipLaunchKernelGGL(k_map_slots, dim3(grid_for((uint64_t)H * 4, 1u << 20)), dim3(256), 0,
                                               st, (const uint32_t*)v0, (const uint32_t*)krf, (uint64_t)H, bk);
                        } else {
                            hipLaunchKernelGGL(k_map_caps, dim3(grid_for((uint64_t)H + 2)), dim3(256), 0, st,
                                               (const uint32_t*)krf, (uint64_t)H, kr);
                            scan_excl(kr, kr, (uint64_t)H + 2, S, S_PRIM, st);
                            if (count)
                                hipLaunchKernelGGL(k_map_fill, dim3(grid_for(count)), dim3(256), 0, st,
                                                   (const uint32_t*)k0, (const uint32_t*)v0, count,
                                                   (const uint32_t*)krf, (const uint32_t*)kr, pos);
                        }
                        c->add_bytes("map_index", (double)count * (16 + 8) + 8.0 * H, (double)count);
                    }
                    unsigned long long* work = S.as<unsigned long long>(S_MWORK, 8);
                    fill<unsigned long long>(c, work, 8, 0ull);
                    if (phased) {
                        Region r(c, phased_region);
                        MapArgs a{hq_pg, ref_len, kr, pos, H - 1, start, packed, W, L, unm, n_unm,
                                  Nascii, nN, last, word, work, P.max_mm, bk};
                        a.nids = Nids;
                        a.nmask = nmask;
                        map_phased(c, a, rc, nm, cnt, learned, trace);
                    } else {
                        Region r(c, "mapping");
                        MapArgs a{hq_pg, ref_len, kr, pos, H - 1, start, packed, W, L, unm, n_unm,
                                  Nascii, nN, last, word, work, P.max_mm, bk};
                        a.nids = Nids;
                        a.nmask = nmask;
                        const dim3 grid(grid_for(nm, c->tune.map_grid));
                        if (bk && W <= 8)
                            launch_map_fast(rc, W, grid, st, a);
                        else if (rc)
                            hipLaunchKernelGGL(k_map<true>, grid, dim3(256), 0, st, a);
                        else
                            hipLaunchKernelGGL(k_map<false>, grid, dim3(256), 0, st, a);
                    }
                    map_stats(phased ? phased_region : "mapping", work);
                }
            }
        }
        // compress.cu:524-594 (get_result)
        uint32_t* rid = S.as<uint32_t>(S_MREC_ID, nm + 1);
        uint64_t* rw = S.as<uint64_t>(S_MREC_W, nm + 1);
        uint8_t* mis = S.as<uint8_t>(S_MIS, nm + 1);
        uint32_t* um = S.as<uint32_t>(S_UNMAPPED, (uint64_t)n_unm + 1);
        uint32_t* umN = S.as<uint32_t>(S_UNMAPPED_N, (uint64_t)nN + 1);
        set_cnt(c, cnt + C_TMP, (uint32_t)nm);
        compact(cnt + C_TMP, PredMapped{last, P.max_mm, 0, true},
                EmitMapRec{last, word, unm, n_unm, rid, rw, mis}, cnt + C_NMAP, S, S_PRIM, st, nm);
        set_cnt(c, cnt + C_TMP, n_unm);
        compact(cnt + C_TMP, PredMapped{last, P.max_mm, 0, false}, EmitUnmapped{unm, um}, cnt + C_NUNMAP, S,
                S_PRIM, st, n_unm);
        set_cnt(c, cnt + C_TMP, nN);
        compact(cnt + C_TMP, PredMapped{last, P.max_mm, n_unm, false}, EmitUnmapped{nullptr, umN},
                cnt + C_NUNMAPN, S, S_PRIM, st, nN);
        const uint32_t nmap = read_cnt(c, cnt + C_NMAP);
        const uint32_t nun = read_cnt(c, cnt + C_NUNMAP);
        const uint32_t nunN = read_cnt(c, cnt + C_NUNMAPN);
        // results go to the host while the LQ stages run (the later stages
        // only read mis, um and umN, at sizes that keep their allocations)
        curc_alignment* drec = S.as<curc_alignment>(S_RECMAP, (uint64_t)nmap + 1);
        if (nmap)
            hipLaunchKernelGGL(k_rec_aos, dim3(grid_for(nmap)), dim3(256), 0, st, (const uint32_t*)rid,
                               (const uint64_t*)rw, (const uint32_t*)nullptr, (uint64_t)nmap, 0u, 0ull, drec);
        // small arrays are pageable and copy synchronously on the compute
        // stream: issue them before the large copies, whose DMA they would
        // otherwise wait behind
        d2h_side(c, B->unmapped_N, (const uint32_t*)umN, nunN, 1);
        d2h_side(c, B->unmapped, (const uint32_t*)um, nun, 1);
        d2h_side(c, B->map_rec, (const curc_alignment*)drec, nmap, 1);
        d2h_side(c, B->map_mis, (const uint8_t*)mis, nmap, 1);
        if (trace) {
            const uint64_t nr = std::min<uint64_t>(read_one(c, S.as<unsigned long long>(S_TRCNT, 1)), c->trace_cap);
            d2h(c, B->trace_x, (const float*)S.as<float>(S_TRX, c->trace_cap * MAP_NFEAT), nr * MAP_NFEAT);
            d2h(c, B->trace_y, (const float*)S.as<float>(S_TRY, c->trace_cap), nr);
            CURC_HIP(stream_wait(st));
        }
    }
    if (stop <= CURC_STAGE_MAP) return;

    // ---- a12 beside a11: in a single-block run the N-read stage runs on a
    // helper context (own stream and arena) in a thread of its own, after
    // the end of mapping, while this thread runs the LQ stage
    const uint32_t n_nlq = (uint32_t)B->unmapped_N.size();
    const uint32_t* umN_dev = S.as<uint32_t>(S_UNMAPPED_N, (uint64_t)nN + 1);
    NlqOut nlq;
    bool nlq_side = false;
    std::exception_ptr nlq_err;
    curc_ctx* hx = nullptr;
    struct Joiner {
        std::thread t;
        ~Joiner() {
            if (t.joinable()) t.join();
        }
    } nlq_thread;
    if (n_nlq && stop > CURC_STAGE_LQ && c->side_copies && c->nev && (hx = helper_ctx(c)) != nullptr) {
        CURC_HIP(hipEventRecord(c->nev, st));
        nlq_side = true;
        const uint32_t items = c->tune.radix_items;
        nlq_thread.t = std::thread([&, items] {
            try {
                CURC_HIP(hipSetDevice(hx->device));
                radix_items() = items;
                nlq = nlq_stage(hx, c->nev, Nascii, nN, umN_dev, n_nlq, L, B);
            } catch (...) {
                nlq_err = std::current_exception();
                // no copy into the block's arrays may outlive an error path
                stream_wait(hx->st);
            }
        });
    }

    // ---- a11: LQ pseudogenome (compress.cu:1649-1722)
    const uint32_t n_lq = (uint32_t)B->unmapped.size();
    if (n_lq) {
        Region r(c, "lq");
        uint64_t* lqp = S.as<uint64_t>(S_LQPACKED, (uint64_t)n_lq * W + 1);
        uint32_t* um = S.as<uint32_t>(S_UNMAPPED, (uint64_t)n_lq + 1);
        hipLaunchKernelGGL(k_gather_packed, dim3(grid_for((uint64_t)n_lq * W)), dim3(256), 0, st,
                           (const uint64_t*)packed, (const uint32_t*)um, (uint64_t)n_lq, W, lqp);
        fill<uint64_t>(c, lqp + (uint64_t)n_lq * W, 1, 0ull);
        uint32_t* lnext = S.as<uint32_t>(S_LQNEXT, n_lq);
        uint32_t* lprev = S.as<uint32_t>(S_LQPREV, n_lq);
        uint16_t* loff = S.as<uint16_t>(S_LQOFF, n_lq);
        fill<uint32_t>(c, lnext, n_lq, INV);
        fill<uint32_t>(c, lprev, n_lq, INV);
        fill<uint16_t>(c, loff, n_lq, INVOFF);
        uint32_t* ids = S.as<uint32_t>(S_PRE, n_lq);
        hipLaunchKernelGGL(k_iota, dim3(grid_for(n_lq)), dim3(256), 0, st, ids, (uint64_t)n_lq);
        set_cnt(c, cnt + C_PRE, n_lq);
        set_cnt(c, cnt + C_SUF, n_lq);
        full_match<Enc2>(c, lqp, W, L, ids, ids, cnt, n_lq, lnext, lprev, loff, "lq");
        if (keep) {
            d2h(c, B->lq_next, lnext, n_lq);
            d2h(c, B->lq_prev, lprev, n_lq);
            d2h(c, B->lq_off, loff, n_lq);
        }
        lq_asm = assemble(c, ids, n_lq, n_lq, lnext, lprev, loff, L, cnt, S_LQIDS);
        const AsmOut& a = lq_asm;
        B->lq_len = a.len;
        uint64_t* lpg = build_pg(c, lqp, W, a, S_PG_LQ);
        records_to_host(c, a, um, 0, B->lq_rec);
        d2h(c, B->lq_pg2, lpg, (a.len + 31) / 32 + 1);
        CURC_HIP(stream_wait(st));
    }
    if (nlq_side) {
        nlq_thread.t.join();
        if (nlq_err) std::rethrow_exception(nlq_err);
        // the helper's stage times go to this context
        CURC_HIP(stream_wait(hx->st));
        hx->harvest();
        for (const StageStat& sst : hx->stats) {
            const int id = c->stat_id(sst.name.c_str());
            c->stats[id].ms += sst.ms;
            c->stats[id].launches += sst.launches;
            c->stats[id].alg_bytes += sst.alg_bytes;
            c->stats[id].items += sst.items;
        }
        hx->stats.clear();
    }
    if (stop <= CURC_STAGE_LQ) return;

    // ---- a12 ran beside a11 (or runs now, on this context)
    if (n_nlq && !nlq_side) nlq = nlq_stage(c, nullptr, Nascii, nN, umN_dev, n_nlq, L, B);
    nlq_asm = nlq.a;

    // ---- a13: record sort + stream builders (compress.cu:1730-2044)
    if (flags & CURC_RUN_STREAMS) {
        Region r(c, "streams");
        build_streams(c, B, hq_asm, lq_asm, nlq_asm, cnt, (flags & CURC_RUN_PRESERVE_ORDER) != 0,
                      (flags & CURC_RUN_PAIRED) != 0);
    }

    // ---- a14: pseudogenome self-match (compress.cu:2053-2069): the LQ and
    // N-read LQ texts (when present) and then the HQ text against the HQ text
    if (flags & CURC_RUN_REFMATCH) {
        Region r(c, "refmatch");
        const uint64_t hq_len = B->hq_len;
        const uint64_t* hq2 = S.as<uint64_t>(S_PG_HQ, (hq_len + 31) / 32 + 1);
        const RmIndex ix = rm_index(c, hq2, hq_len);
        const bool keepm = keep;
        if (B->lq_len)
            rm_match(c, ix, hq2, hq_len, S.as<uint64_t>(S_PG_LQ, (B->lq_len + 31) / 32 + 1), nullptr, B->lq_len,
                     false, keepm, cnt, B->rm[1]);
        if (!B->nlq_text.empty())
            rm_match(c, ix, hq2, hq_len, nullptr, nlq.npg ? nlq.npg : S.as<uint8_t>(S_NPG, B->nlq_text.size() + 16),
                     B->nlq_text.size(), false, keepm, cnt, B->rm[2]);
        rm_match(c, ix, hq2, hq_len, hq2, nullptr, hq_len, true, keepm, cnt, B->rm[0]);
    }
}

int fail(curc_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

// One block from ASCII reads (ascii, n_reads) or from FASTQ bytes on the
// device (fq); *out receives the block.
int block_run(curc_ctx* c, const char* ascii, uint64_t n_reads, uint32_t L, int stop_after, uint32_t flags,
              const FqInput* fq, curc_block** out) {
    curc_params P;
    if (curc_make_params(L, &P) != CURC_OK) return fail(c, CURC_EINVAL, "unsupported read length");
    curc_block* B = new curc_block();
    bool done_recorded = false;
    auto finish_input = [&]() {
        // the input's consumer event is recorded on every path, so an uploader waiting on it proceeds
        if (fq && fq->done && !done_recorded) hipEventRecord(fq->done, c->st);
        done_recorded = true;
    };
    try {
        CURC_HIP(hipSetDevice(c->device));
        radix_items() = c->tune.radix_items;
        run_block(c, ascii, n_reads, L, stop_after <= 0 ? CURC_STAGE_ALL : stop_after, flags, B, fq);
        done_recorded = true;
        CURC_HIP(hipGetLastError());
        CURC_HIP(stream_wait(c->st));
        join_side(c);
        c->harvest();
        if (flags & CURC_RUN_KEEP_STAGES) {
            // expose the packed reads and N-read text for checking
            const uint64_t N = B->n_reads;
            d2h(c, B->packed, c->S.as<uint64_t>(S_PACKED, N * P.W + 1), N * P.W);
            std::vector<uint8_t> na;
            d2h(c, na, c->S.as<uint8_t>(S_NASCII, (uint64_t)B->Nids.size() * L + 16), (uint64_t)B->Nids.size() * L);
            CURC_HIP(stream_wait(c->st));
            B->Nascii.assign(na.begin(), na.end());
        }
    } catch (const std::exception& e) {
        // copies into B's arrays may be in flight: join them before freeing
        finish_input();
        stream_wait(c->st);
        if (c->cst) stream_wait(c->cst);
        delete B;
        c->pending.clear();
        std::string m = e.what();
        int code = CURC_EHIP;
        if (dynamic_cast<const InputError*>(&e)) code = CURC_EINVAL;
        else if (dynamic_cast<const std::length_error*>(&e)) code = CURC_ELIMIT;
        else if (m.find("out of memory") != std::string::npos) code = CURC_ENOMEM;
        return fail(c, code, m);
    }
    *out = B;
    return CURC_OK;
}

// Runs n_jobs jobs on up to `workers` contexts (c and its worker contexts,
// each with its own stream and scratch arena), heaviest first by `weight`
// (stable), each by run(ctx, job, &out[job]).  Worker errors and stage
// times land on c.
template <class Fn>
int run_on_workers(curc_ctx* c, uint32_t n_jobs, uint32_t workers, const std::vector<uint64_t>& weight,
                   curc_block** out, int* status, Fn run) {
    for (uint32_t j = 0; j < n_jobs; ++j) {
        out[j] = nullptr;
        if (status) status[j] = CURC_OK;
    }
    if (n_jobs == 0) return CURC_OK;
    if (workers == 0) workers = c->tune.workers;
    workers = std::min<uint32_t>(std::min<uint32_t>(workers, CURC_BATCH_MAX_WORKERS), n_jobs);
    while (c->workers.size() + 1 < workers) {
        curc_ctx* w = curc_ctx_create(c->device);
        if (!w) return fail(c, CURC_EHIP, "could not create a worker stream");
        w->side_copies = false;
        c->workers.push_back(w);
    }
    // the batch's contexts copy results on their own streams (one stream
    // per context); the caller's copy stream is released for the batch
    hipSetDevice(c->device);
    if (c->cst) {
        stream_wait(c->cst);
        hipStreamDestroy(c->cst);
        c->cst = nullptr;
    }
    const bool side = c->side_copies;
    c->side_copies = false;
    std::vector<curc_ctx*> pool{c};
    for (uint32_t w = 1; w < workers; ++w) {
        curc_ctx* wc = c->workers[w - 1];
        wc->timing = c->timing;
        wc->index_csr = c->index_csr;
        wc->tune = c->tune;
        wc->graphs.enabled = c->graphs.enabled;
        // workers share the scorer's device arrays (read-only during mapping)
        wc->has_scorer = c->has_scorer;
        wc->scorer_net = c->scorer_net;
        wc->scorer_threshold = c->scorer_threshold;
        wc->scorer_threshold_unmapped = c->scorer_threshold_unmapped;
        wc->trace_every = c->trace_every;
        wc->trace_cap = c->trace_cap;
        pool.push_back(wc);
    }
    std::vector<uint32_t> order(n_jobs);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return weight[a] > weight[b]; });
    std::atomic<uint32_t> next{0};
    std::vector<int> rc(n_jobs, CURC_OK);
    auto work = [&](curc_ctx* wc) {
        hipSetDevice(wc->device);
        for (uint32_t k = next++; k < n_jobs; k = next++) {
            const uint32_t j = order[k];
            rc[j] = run(wc, j, &out[j]);
        }
    };
    std::vector<std::thread> th;
    for (uint32_t w = 1; w < workers; ++w) th.emplace_back(work, pool[w]);
    work(c);
    for (auto& t : th) t.join();
    int first = CURC_OK;
    for (uint32_t j = 0; j < n_jobs; ++j) {
        if (status) status[j] = rc[j];
        if (rc[j] != CURC_OK && first == CURC_OK) first = rc[j];
    }
    for (uint32_t w = 1; w < workers; ++w) {
        curc_ctx* wc = pool[w];
        if (first != CURC_OK && c->err.empty()) c->err = wc->err;
        for (const StageStat& st : wc->stats) {
            const int id = c->stat_id(st.name.c_str());
            c->stats[id].ms += st.ms;
            c->stats[id].launches += st.launches;
            c->stats[id].alg_bytes += st.alg_bytes;
            c->stats[id].items += st.items;
        }
        wc->stats.clear();
        c->graphs.eager += wc->graphs.eager;
        c->graphs.captures += wc->graphs.captures;
        c->graphs.replays += wc->graphs.replays;
        wc->graphs.eager = wc->graphs.captures = wc->graphs.replays = 0;
    }
    c->side_copies = side;
    return first;
}

// Copies a FASTQ job's bytes into the device buffer `dev` (file 2 at the
// next 256-byte boundary after file 1) on stream s; returns the layout.
FqInput fq_upload(const curc_fastq_job& j, uint8_t* dev, hipStream_t s) {
    FqInput in;
    const uint64_t off2 = (j.bytes1 + 255) & ~255ull;
    in.fq[0] = dev;
    in.bytes[0] = j.bytes1;
    if (j.bytes1) CURC_HIP(hipMemcpyAsync(dev, j.fq1, j.bytes1, hipMemcpyHostToDevice, s));
    if (j.fq2) {
        in.fq[1] = dev + off2;
        in.bytes[1] = j.bytes2;
        if (j.bytes2) CURC_HIP(hipMemcpyAsync(dev + off2, j.fq2, j.bytes2, hipMemcpyHostToDevice, s));
    }
    return in;
}

uint64_t fq_dev_bytes(const curc_fastq_job& j) {
    return ((j.bytes1 + 255) & ~255ull) + (j.fq2 ? j.bytes2 : 0) + 256;
}

// Rows a host-parsed file of a paired job may hold (a record takes at least
// L + 4 bytes): the stride of each file's region in the device buffer.
uint64_t fq_pair_rows(const curc_fastq_job& j) { return std::max(j.bytes1, j.bytes2) / ((uint64_t)j.L + 4) + 2; }

// Device bytes a job needs: its FASTQ bytes, or with host parsing its
// sequences (at most the FASTQ bytes: a record takes at least L + 4 bytes);
// a paired block holds the interleaved rows, then each file's rows.
uint64_t fq_dev_bytes(const curc_fastq_job& j, uint32_t host_threads) {
    if (!host_threads || !j.fq2) return fq_dev_bytes(j);
    return std::max(fq_dev_bytes(j), 4 * fq_pair_rows(j) * j.L + 512);
}

// Host threads for a file of nb bytes: ranges of at least 1 MB each.
uint32_t fq_host_threads(uint32_t want, uint64_t nb) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, nb >> 20));
}

// Rows 2k (file 1's record k) and 2k+1 (file 2's) of a paired block.
__global__ void k_interleave_rows(const uint8_t* __restrict__ a, const uint8_t* __restrict__ b, uint64_t n1,
                                  uint32_t L, uint8_t* __restrict__ out) {
    GRID_STRIDE(i, 2 * n1 * L) {
        const uint64_t r = i / L, j = i - r * L;
        out[i] = ((r & 1) ? b : a)[(r >> 1) * L + j];
    }
}

// Host parse of a job's FASTQ bytes (fq_host_parse, `threads` threads) into
// the pinned buffer c->stage(which), with each parsed range's sequences
// copied to the device buffer `dev` (fq_dev_bytes(j, threads) bytes) on
// stream s as soon as it is ready, while the later ranges are parsed.  The
// caller makes sure earlier copies out of that pinned buffer have ended.
// Format errors throw InputError with the device parser's messages, in its
// order (preprocess.cpp:147-199).
FqInput fq_stage_host(curc_ctx* c, const curc_fastq_job& j, uint32_t threads, int which, uint8_t* dev,
                      hipStream_t s) {
    const int nf = j.fq2 ? 2 : 1;
    const uint64_t nb[2] = {j.bytes1, j.fq2 ? j.bytes2 : 0};
    const uint8_t* src[2] = {(const uint8_t*)j.fq1, (const uint8_t*)j.fq2};
    uint32_t th[2];
    uint64_t cap[2] = {0, 0};
    for (int f = 0; f < nf; ++f) {
        th[f] = fq_host_threads(threads, nb[f]);
        cap[f] = fq_host_cap(nb[f], j.L, th[f]);
    }
    uint8_t* h = c->stage(which, cap[0] + cap[1] + 64);
    const uint64_t L = j.L;
    // single-end rows straight to their place; a paired block's files to
    // regions after the interleaved rows, interleaved on the device below
    const uint64_t pr = nf == 2 ? fq_pair_rows(j) : 0;
    FqHostParse r[2];
    for (int f = 0; f < nf; ++f) {
        uint8_t* d = nf == 1 ? dev : dev + (2 + f) * pr * L;
        const uint8_t* hf = h + (f ? cap[0] : 0);
        fq_host_parse(src[f], nb[f], j.L, th[f], h + (f ? cap[0] : 0), &r[f], [&](const FqSeg& g, uint64_t first) {
            CURC_HIP(hipMemcpyAsync(d + first * L, hf + g.off, g.count * L, hipMemcpyHostToDevice, s));
        });
    }
    for (int f = 0; f < nf; ++f)
        if (r[f].err == FQH_START) throw InputError("FASTQ block does not start with '@'");
    for (int f = 0; f < nf; ++f)
        if (r[f].err == FQH_TRUNC) throw InputError("truncated FASTQ record");
    if (nf == 2 && r[0].n != r[1].n) throw InputError("paired-end files hold different record counts");
    for (int f = 0; f < nf; ++f)
        if (r[f].err == FQH_LEN) throw InputError("read length isn't fix");  // preprocess.cpp:154
    const uint64_t n1 = r[0].n, N = n1 * nf;
    if (N >= (uint64_t)INV) throw std::length_error("a block holds at most 2^32-2 reads");
    FqInput in;
    in.n_ascii = N;
    in.ascii = dev;
    if (nf == 2 && n1)
        hipLaunchKernelGGL(k_interleave_rows, dim3(grid_for(2 * n1 * L)), dim3(256), 0, s,
                           (const uint8_t*)(dev + 2 * pr * L), (const uint8_t*)(dev + 3 * pr * L), n1, j.L, dev);
    return in;
}

int fq_check_job(curc_ctx* c, const curc_fastq_job* j) {
    if (!j) return fail(c, CURC_EINVAL, "null job");
    if ((j->bytes1 && !j->fq1) || (j->fq2 == nullptr && j->bytes2)) return fail(c, CURC_EINVAL, "null input");
    curc_params P;
    if (curc_make_params(j->L, &P) != CURC_OK) return fail(c, CURC_EINVAL, "unsupported read length");
    return CURC_OK;
}

uint32_t fq_flags(const curc_fastq_job& j) { return j.flags | (j.fq2 ? CURC_RUN_PAIRED : 0u); }

// One FASTQ job: bytes already on the device, or copied into scratch slot
// `slot` first on the context's stream.
int fq_block_run(curc_ctx* c, const curc_fastq_job& j, int stop_after, int slot, curc_block** out,
                 uint32_t host_threads) {
    int rc = fq_check_job(c, &j);
    if (rc != CURC_OK) return rc;
    FqInput in;
    try {
        CURC_HIP(hipSetDevice(c->device));
        if (j.flags & CURC_RUN_INPUT_ON_DEVICE) {
            in.fq[0] = (const uint8_t*)j.fq1;
            in.bytes[0] = j.bytes1;
            in.fq[1] = (const uint8_t*)j.fq2;
            in.bytes[1] = j.fq2 ? j.bytes2 : 0;
        } else if (host_threads) {
            // the previous block's copies out of the pinned buffer ended with
            // its block (block_run waits for the stream)
            in = fq_stage_host(c, j, host_threads, 0, c->S.as<uint8_t>(slot, fq_dev_bytes(j, host_threads)), c->st);
        } else {
            in = fq_upload(j, c->S.as<uint8_t>(slot, fq_dev_bytes(j)), c->st);
        }
    } catch (const std::exception& e) {
        stream_wait(c->st);
        return fail(c, dynamic_cast<const InputError*>(&e) ? CURC_EINVAL : CURC_EHIP, e.what());
    }
    return block_run(c, nullptr, 0, j.L, stop_after, fq_flags(j), &in, out);
}

}  // namespace

namespace {

// Device copy of a host array (scratch slot `slot`), or nullptr when empty.
template <typename T>
T* dec_upload(curc_ctx* c, int slot, const T* src, uint64_t n) {
    T* d = c->S.as<T>(slot, n + 1);
    if (n) h2d(c, d, src, n);
    return d;
}

// Read reconstruction of one block (decompress.cu:909-1005, :472-577) from
// host streams into host output lines.
void decode_run(curc_ctx* c, const curc_decode_job& j, char* out1, char* out2) {
    Scratch& S = c->S;
    hipStream_t st = c->st;
    const bool preserve = (j.flags & CURC_RUN_PRESERVE_ORDER) != 0;
    const bool paired = (j.flags & CURC_RUN_PAIRED) != 0;
    const uint32_t L = j.L, LL = L + 1;
    const uint64_t n = j.n_reads;
    if (paired && (n & 1)) throw InputError("a paired block holds an even number of reads");
    if (!preserve && j.n_rec + j.n_lq + j.n_nlq != n) throw InputError("record counts do not add up to the reads");
    uint64_t n_mis = 0;
    for (uint64_t r = 0; r < j.n_rec; ++r) n_mis += j.mis_count[r];
    if (n_mis != j.n_mis) throw InputError("mismatch symbols do not match the mismatch counts");
    uint32_t* cnt = S.as<uint32_t>(S_CNT, C_COUNT);
    fill<uint32_t>(c, cnt, C_COUNT, 0u);
    DecArgs a{};
    a.L = L;
    a.paired = paired ? 1u : 0u;
    a.n_reads = n;
    a.hq_len = j.hq_len;
    a.lq_len = j.lq_len;
    a.nlq_len = j.nlq_len;
    a.n_mis = n_mis;
    a.err = cnt + C_FQERR;
    {
        Region r(c, "decode_h2d");
        uint8_t* text = S.as<uint8_t>(S_D_TEXT, j.hq_len + j.lq_len + j.nlq_len + 1);
        if (j.hq_len) h2d(c, text, (const uint8_t*)j.hq, j.hq_len);
        if (j.lq_len) h2d(c, text + j.hq_len, (const uint8_t*)j.lq, j.lq_len);
        if (j.nlq_len) h2d(c, text + j.hq_len + j.lq_len, (const uint8_t*)j.nlq, j.nlq_len);
        a.text = text;
        a.strand = (const char*)dec_upload(c, S_D_STRAND, (const uint8_t*)j.strand, j.n_rec);
        a.mis_count = dec_upload(c, S_D_CNT, j.mis_count, j.n_rec);
        a.mis_base = dec_upload(c, S_D_BASE, j.mis_base, n_mis);
        a.table = dec_upload(c, S_D_TAB, j.mis_table, 16);
        if (preserve) {
            a.id_pos = dec_upload(c, S_D_IDPOS, j.id_pos, n);
        } else {
            a.pe_slot = paired ? dec_upload(c, S_D_PE, j.pe_slot, n) : nullptr;
        }
    }
    // offset streams, one after another; sstart[m] = first u16 of stream m
    std::vector<uint64_t> sstart(j.max_mm + 2, 0);
    for (uint32_t m = 1; m <= j.max_mm; ++m) sstart[m + 1] = sstart[m] + j.mis_off_len[m - 1];
    uint16_t* streams = S.as<uint16_t>(S_D_STREAMS, sstart[j.max_mm + 1] + 1);
    {
        Region r(c, "decode_h2d");
        for (uint32_t m = 1; m <= j.max_mm; ++m)
            if (j.mis_off_len[m - 1]) h2d(c, streams + sstart[m], j.mis_off[m - 1], j.mis_off_len[m - 1]);
    }
    uint64_t* d_sstart = dec_upload(c, S_D_SSTART, sstart.data(), sstart.size());
    const uint16_t* read_off = nullptr;
    const uint16_t* lq_off = nullptr;
    const uint16_t* nlq_off = nullptr;
    if (!preserve) {
        Region r(c, "decode_h2d");
        read_off = dec_upload(c, S_D_ROFF, j.read_off, j.n_rec);
        lq_off = dec_upload(c, S_D_LQ, j.lq_off, j.n_lq);
        nlq_off = dec_upload(c, S_D_NLQ, j.nlq_off, j.n_nlq);
    }
    uint8_t* o1 = S.as<uint8_t>(S_D_OUT1, (paired ? n / 2 : n) * LL + 1);
    uint8_t* o2 = paired ? S.as<uint8_t>(S_D_OUT2, n / 2 * LL + 1) : nullptr;
    a.out1 = o1;
    a.out2 = o2;
    uint64_t hq_reads = 0;
    {
        Region r(c, "decode");
        // positions (order not preserved) or HQ record indexes (preserved)
        if (!preserve) {
            uint64_t* pv = S.as<uint64_t>(S_D_PV, n + 1);
            uint64_t* px = S.as<uint64_t>(S_D_PX, n + 1);
            if (n)
                hipLaunchKernelGGL(k_dec_pos_values, dim3(grid_for(n)), dim3(256), 0, st, read_off, j.n_rec, lq_off,
                                   j.n_lq, nlq_off, j.n_nlq, L, pv);
            scan_excl_t<uint64_t>(pv, px, j.n_rec, S, S_PRIM, st);
            scan_excl_t<uint64_t>(pv + j.n_rec, px + j.n_rec, j.n_lq, S, S_PRIM, st);
            scan_excl_t<uint64_t>(pv + j.n_rec + j.n_lq, px + j.n_rec + j.n_lq, j.n_nlq, S, S_PRIM, st);
            a.pos_v = pv;
            a.pos_ex = px;
            a.nh = j.n_rec;
            a.nl = j.n_lq;
        } else {
            uint64_t* hv = S.as<uint64_t>(S_D_HQV, n + 1);
            uint64_t* hx = S.as<uint64_t>(S_D_HQX, n + 1);
            if (n) hipLaunchKernelGGL(k_dec_hq_mask, dim3(grid_for(n)), dim3(256), 0, st, a.id_pos, n, j.hq_len, hv);
            scan_excl_t<uint64_t>(hv, hx, n, S, S_PRIM, st);
            a.hq_idx = hx;
            uint64_t last[2] = {0, 0};
            if (n) {
                CURC_HIP(hipMemcpyAsync(&last[0], hx + n - 1, 8, hipMemcpyDeviceToHost, st));
                CURC_HIP(hipMemcpyAsync(&last[1], hv + n - 1, 8, hipMemcpyDeviceToHost, st));
                CURC_HIP(stream_wait(st));
            }
            hq_reads = last[0] + last[1];
            if (hq_reads != j.n_rec) throw InputError("reads placed in the HQ text do not match the records");
        }
        // first mismatch of each record
        uint64_t* bv = S.as<uint64_t>(S_D_BV, j.n_rec + 1);
        uint64_t* bx = S.as<uint64_t>(S_D_BX, j.n_rec + 1);
        if (j.n_rec)
            hipLaunchKernelGGL(k_dec_u8_to_u64, dim3(grid_for(j.n_rec)), dim3(256), 0, st, a.mis_count, j.n_rec, bv);
        scan_excl_t<uint64_t>(bv, bx, j.n_rec, S, S_PRIM, st);
        a.bidx = bx;
        // mismatch offsets in record order
        uint16_t* flat = S.as<uint16_t>(S_D_FLAT, n_mis + 1);
        if (j.n_rec && n_mis) {
            uint32_t* k0 = S.as<uint32_t>(S_D_K0, j.n_rec + 1);
            uint32_t* k1 = S.as<uint32_t>(S_D_K1, j.n_rec + 1);
            uint32_t* v0 = S.as<uint32_t>(S_D_V0, j.n_rec + 1);
            uint32_t* v1 = S.as<uint32_t>(S_D_V1, j.n_rec + 1);
            uint64_t* gs = S.as<uint64_t>(S_D_GS, 257);
            hipLaunchKernelGGL(k_dec_count_keys, dim3(grid_for(j.n_rec)), dim3(256), 0, st, a.mis_count, j.n_rec, k0,
                               v0);
            radix_sort_pairs<uint32_t>(k0, v0, k1, v1, j.n_rec, 0, 8, S, S_PRIM, st);
            hipLaunchKernelGGL(k_dec_group_start, dim3(grid_for(j.n_rec)), dim3(256), 0, st, (const uint32_t*)k0,
                               j.n_rec, gs);
            hipLaunchKernelGGL(k_dec_offsets, dim3(grid_for(j.n_rec)), dim3(256), 0, st, (const uint32_t*)k0,
                               (const uint32_t*)v0, j.n_rec, (const uint64_t*)gs, (const uint16_t*)streams,
                               (const uint64_t*)d_sstart, (const uint64_t*)bx, L, j.max_mm, flat, a.err);
        }
        a.flat_off = flat;
        const uint64_t wgs = (n + DEC_LINES - 1) / DEC_LINES;
        if (wgs) {
            if (preserve)
                hipLaunchKernelGGL(k_dec_reads<true>, dim3((uint32_t)wgs), dim3(256), DEC_LINES * LL, st, a);
            else
                hipLaunchKernelGGL(k_dec_reads<false>, dim3((uint32_t)wgs), dim3(256), DEC_LINES * LL, st, a);
        }
        c->add_bytes("decode", (double)n * (L + LL) + (double)j.n_rec * 18 + (double)n_mis * 5,
                     (double)n);
    }
    uint32_t e = 0;
    CURC_HIP(hipMemcpyAsync(&e, a.err, 4, hipMemcpyDeviceToHost, st));
    CURC_HIP(stream_wait(st));
    if (e) throw InputError("corrupt streams (offsets or positions out of range)");
    {
        Region r(c, "decode_d2h");
        const uint64_t lines1 = paired ? n / 2 : n;
        if (lines1) CURC_HIP(hipMemcpyAsync(out1, o1, lines1 * LL, hipMemcpyDeviceToHost, st));
        if (paired && lines1) CURC_HIP(hipMemcpyAsync(out2, o2, lines1 * LL, hipMemcpyDeviceToHost, st));
    }
    CURC_HIP(stream_wait(st));
}

}  // namespace

extern "C" {

int curc_make_params(uint32_t L, curc_params* p) {
    // preprocess.cpp:296-318
    if (!p || L == 0 || L > 512) return CURC_EINVAL;
    p->L = L;
    p->W = (L + 31) / 32;
    p->max_off = L * 65 / 100;
    p->k = (L < 89) ? 36 * L / 100 : 32;
    if (p->k == 0 || p->k + p->max_off > L) return CURC_EINVAL;
    p->kip = L - p->k - p->max_off;
    p->max_mm = L / 5;
    return CURC_OK;
}

int curc_device_count(int* count) {
    if (!count) return CURC_EINVAL;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return CURC_EHIP;
    }
    *count = n;
    return CURC_OK;
}

curc_ctx* curc_ctx_create(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    // the fixed scratch of curc_fastq_pack_static on this device
    if (!static_scratch_init(device)) return nullptr;
    curc_ctx* c = new curc_ctx();
    c->device = device;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            c->cus = (uint32_t)prop.multiProcessorCount;
    }
    const char* ix = std::getenv("CURC_INDEX");
    c->index_csr = ix && std::string(ix) == "csr";
    if (const char* t = std::getenv("CURC_TUNE")) parse_tuning(t, c->tune);
    const char* gr = std::getenv("CURC_GRAPHS");
    c->graphs.enabled = gr && std::string(gr) == "1";
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->cev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->cev[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->nev, hipEventDisableTiming) != hipSuccess) {
        if (c->nev) hipEventDestroy(c->nev);
        if (c->cev[1]) hipEventDestroy(c->cev[1]);
        if (c->cev[0]) hipEventDestroy(c->cev[0]);
        if (c->st) hipStreamDestroy(c->st);
        delete c;
        return nullptr;
    }
    return c;
}

void curc_ctx_destroy(curc_ctx* c) {
    if (!c) return;
    for (curc_ctx* w : c->workers) curc_ctx_destroy(w);
    c->workers.clear();
    if (c->helper) curc_ctx_destroy(c->helper);
    c->helper = nullptr;
    hipSetDevice(c->device);
    stream_wait(c->st);
    if (c->cst) stream_wait(c->cst);
    c->graphs.clear();
    c->S.release();
    for (auto& p : c->pending) {
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    for (auto e : c->event_pool) hipEventDestroy(e);
    hipEventDestroy(c->cev[0]);
    hipEventDestroy(c->cev[1]);
    hipEventDestroy(c->nev);
    for (auto e : c->mev)
        if (e) hipEventDestroy(e);
    if (c->cst) hipStreamDestroy(c->cst);
    if (c->ust) {
        stream_wait(c->ust);
        hipStreamDestroy(c->ust);
    }
    for (auto& h : c->hstage)
        if (h) hipHostFree(h);
    hipStreamDestroy(c->st);
    delete c;
}

const char* curc_ctx_last_error(curc_ctx* c) { return c ? c->err.c_str() : "null context"; }

int curc_ctx_trim(curc_ctx* c) {
    if (!c) return CURC_EINVAL;
    hipSetDevice(c->device);
    stream_wait(c->st);
    if (c->cst) stream_wait(c->cst);
    c->graphs.clear();
    c->S.release();
    for (curc_ctx* w : c->workers) curc_ctx_trim(w);
    if (c->helper) curc_ctx_trim(c->helper);
    HostPool::get().trim();
    return CURC_OK;
}

int curc_host_alloc(curc_ctx* c, void** p, uint64_t bytes) {
    if (!c || !p) return CURC_EINVAL;
    hipSetDevice(c->device);
    if (hipHostMalloc(p, bytes, hipHostMallocPortable) != hipSuccess) return fail(c, CURC_ENOMEM, "hipHostMalloc failed");
    return CURC_OK;
}
int curc_host_free(curc_ctx* c, void* p) {
    if (!c) return CURC_EINVAL;
    return hipHostFree(p) == hipSuccess ? CURC_OK : fail(c, CURC_EHIP, "hipHostFree failed");
}
int curc_dev_alloc(curc_ctx* c, void** p, uint64_t bytes) {
    if (!c || !p) return CURC_EINVAL;
    hipSetDevice(c->device);
    if (hipMalloc(p, bytes) != hipSuccess) return fail(c, CURC_ENOMEM, "hipMalloc failed");
    return CURC_OK;
}
int curc_dev_free(curc_ctx* c, void* p) {
    if (!c) return CURC_EINVAL;
    hipSetDevice(c->device);
    return hipFree(p) == hipSuccess ? CURC_OK : fail(c, CURC_EHIP, "hipFree failed");
}
int curc_memcpy_h2d(curc_ctx* c, void* d, const void* s, uint64_t bytes) {
    if (!c) return CURC_EINVAL;
    hipSetDevice(c->device);
    if (hipMemcpyAsync(d, s, bytes, hipMemcpyHostToDevice, c->st) != hipSuccess ||
        stream_wait(c->st) != hipSuccess)
        return fail(c, CURC_EHIP, "host to device copy failed");
    return CURC_OK;
}
int curc_memcpy_d2h(curc_ctx* c, void* d, const void* s, uint64_t bytes) {
    if (!c) return CURC_EINVAL;
    hipSetDevice(c->device);
    if (hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
        stream_wait(c->st) != hipSuccess)
        return fail(c, CURC_EHIP, "device to host copy failed");
    return CURC_OK;
}
int curc_stream_copy(curc_ctx* c, void* dst, const void* src, uint64_t bytes, int reps, float* ms) {
    if (!c || !dst || !src || reps < 1 || (bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
        return CURC_EINVAL;
    try {
        CURC_HIP(hipSetDevice(c->device));
        const uint64_t n = bytes / 16;
        const dim3 grid(grid_for(n, 1u << 16));
        hipLaunchKernelGGL(k_stream_copy, grid, dim3(256), 0, c->st, (const uint4*)src, (uint4*)dst, n);  // warm
        hipEvent_t a = c->ev(), b = c->ev();
        CURC_HIP(hipEventRecord(a, c->st));
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_stream_copy, grid, dim3(256), 0, c->st, (const uint4*)src, (uint4*)dst, n);
        CURC_HIP(hipEventRecord(b, c->st));
        CURC_HIP(hipEventSynchronize(b));
        float t = 0;
        CURC_HIP(hipEventElapsedTime(&t, a, b));
        c->event_pool.push_back(a);
        c->event_pool.push_back(b);
        if (ms) *ms = t / (float)reps;
    } catch (const std::exception& e) {
        return fail(c, CURC_EHIP, e.what());
    }
    return CURC_OK;
}

int curc_sync(curc_ctx* c) {
    if (!c) return CURC_EINVAL;
    hipSetDevice(c->device);
    return stream_wait(c->st) == hipSuccess ? CURC_OK : fail(c, CURC_EHIP, "stream sync failed");
}

int curc_device_sync(curc_ctx* c) {
    if (!c) return CURC_EINVAL;
    hipSetDevice(c->device);
    return hipDeviceSynchronize() == hipSuccess ? CURC_OK : fail(c, CURC_EHIP, "device sync failed");
}

int curc_pack(curc_ctx* c, const char* d_ascii, uint64_t n, uint32_t L, uint64_t* d_packed, uint8_t* d_isN) {
    if (!c || (!d_ascii && n) || (!d_packed && n) || L == 0 || L > 512) return CURC_EINVAL;
    try {
        CURC_HIP(hipSetDevice(c->device));
        const uint32_t W = (L + 31) / 32;
        if (n) {
            if (d_isN) fill<uint8_t>(c, d_isN, n, (uint8_t)0);
            hipLaunchKernelGGL(k_pack_w, dim3(grid_for(n * W, 1u << 20)), dim3(256), 0, c->st,
                               (const uint8_t*)d_ascii, n, L, W, d_packed, d_isN);
            CURC_HIP(hipGetLastError());
        }
        CURC_HIP(stream_wait(c->st));
    } catch (const std::exception& e) {
        return fail(c, CURC_EHIP, e.what());
    }
    return CURC_OK;
}

int curc_block_run(curc_ctx* c, const char* ascii, uint64_t n_reads, uint32_t L, int stop_after, uint32_t flags,
                   curc_block** out) {
    if (!c || !out) return CURC_EINVAL;
    *out = nullptr;
    if (n_reads >= (uint64_t)INV) return fail(c, CURC_ELIMIT, "a block holds at most 2^32-2 reads");
    if (n_reads && !ascii) return fail(c, CURC_EINVAL, "null input");
    return block_run(c, ascii, n_reads, L, stop_after, flags, nullptr, out);
}

uint64_t curc_block_get(curc_block* B, int field, const void** ptr) {
    if (!B || !ptr) return 0;
    auto put = [&](const auto& v) -> uint64_t {
        *ptr = v.data();
        return v.size();
    };
    switch (field) {
        case CURC_F_PACKED: return put(B->packed);
        case CURC_F_NONN_IDS: return put(B->nonN);
        case CURC_F_N_IDS: return put(B->Nids);
        case CURC_F_SORTED_IDS: return put(B->sorted);
        case CURC_F_NEXT_DEDUP: return put(B->next_dedup);
        case CURC_F_PREV_DEDUP: return put(B->prev_dedup);
        case CURC_F_NEXT_OVERLAP: return put(B->next_ov);
        case CURC_F_PREV_OVERLAP: return put(B->prev_ov);
        case CURC_F_OFFSET_OVERLAP: return put(B->off_ov);
        case CURC_F_NEXT_PRUNE: return put(B->next_pr);
        case CURC_F_PREV_PRUNE: return put(B->prev_pr);
        case CURC_F_NEXT: return put(B->next);
        case CURC_F_PREV: return put(B->prev);
        case CURC_F_OFFSET: return put(B->offset);
        case CURC_F_HQ_IDS: return put(B->hq);
        case CURC_F_UNMATCHED_IDS: return put(B->unmatched);
        case CURC_F_HQ_PG:
            if (!B->hq_text_ok) {
                B->hq_text = B->hq_pg2.empty() ? std::string() : decode_pg(B->hq_pg2, B->hq_len);
                B->hq_text_ok = true;
            }
            return put(B->hq_text);
        case CURC_F_LQ_PG:
            if (!B->lq_text_ok) {
                B->lq_text = B->lq_pg2.empty() ? std::string() : decode_pg(B->lq_pg2, B->lq_len);
                B->lq_text_ok = true;
            }
            return put(B->lq_text);
        case CURC_F_HQ_PG2: return put(B->hq_pg2);
        case CURC_F_HQ_REC: return put(B->hq_rec);
        case CURC_F_MAP_REC: return put(B->map_rec);
        case CURC_F_MAP_MIS: return put(B->map_mis);
        case CURC_F_UNMAPPED: return put(B->unmapped);
        case CURC_F_UNMAPPED_N: return put(B->unmapped_N);
        case CURC_F_LQ_REC: return put(B->lq_rec);
        case CURC_F_LQ_NEXT: return put(B->lq_next);
        case CURC_F_LQ_PREV: return put(B->lq_prev);
        case CURC_F_LQ_OFFSET: return put(B->lq_off);
        case CURC_F_N_ASCII: return put(B->Nascii);
        case CURC_F_NLQ_PG: return put(B->nlq_text);
        case CURC_F_NLQ_REC: return put(B->nlq_rec);
        case CURC_F_S_STRAND: return put(B->s_strand);
        case CURC_F_S_READ_OFF: return put(B->s_read_off);
        case CURC_F_S_MIS_COUNT: return put(B->s_mis_count);
        case CURC_F_S_MIS_BASE: return put(B->s_mis_base);
        case CURC_F_S_MIS_TABLE: return put(B->s_mis_table);
        case CURC_F_S_LQ_OFF: return put(B->s_lq_off);
        case CURC_F_S_NLQ_OFF: return put(B->s_nlq_off);
        case CURC_F_S_PE_ORDER: return put(B->s_pe_order);
        case CURC_F_S_ID_POS: return put(B->s_id_pos);
        case CURC_F_S_SORTED_REC: return put(B->s_sorted);
        case CURC_F_TRACE_X: *ptr = B->trace_x.data(); return B->trace_x.size() / MAP_NFEAT;
        case CURC_F_TRACE_Y: return put(B->trace_y);
        default:
            if (field >= CURC_F_RM_BASE && field < CURC_F_RM_BASE + 15) {
                const int p = (field - CURC_F_RM_BASE) / 5, k = (field - CURC_F_RM_BASE) % 5;
                const auto& R = B->rm[p];
                switch (k) {
                    case 0: *ptr = R.text.data(); return R.text.size();
                    case 1: *ptr = R.map_off.data(); return R.map_off.size();
                    case 2: *ptr = R.map_len.data(); return R.map_len.size();
                    case 3: *ptr = R.matches.data(); return R.matches.size() / 3;
                    default: *ptr = R.found.data(); return R.found.size() / 3;
                }
            }
            if (field > CURC_F_S_MIS_OFF && (size_t)(field - CURC_F_S_MIS_OFF) + 1 < B->s_mis_off_start.size()) {
                const size_t m = (size_t)(field - CURC_F_S_MIS_OFF);
                const uint64_t a = B->s_mis_off_start[m], b = B->s_mis_off_start[m + 1];
                *ptr = B->s_mis_off.data() + a;
                return b - a;
            }
            *ptr = nullptr;
            return 0;
    }
}

void curc_block_free(curc_block* B) { delete B; }

int curc_ctx_set_timing(curc_ctx* c, int on) {
    if (!c) return CURC_EINVAL;
    c->timing = on != 0;
    return CURC_OK;
}

int curc_ctx_stats(curc_ctx* c, curc_stage_stat* out, int max) {
    if (!c) return CURC_EINVAL;
    std::vector<StageStat> all = c->stats;
    // offset loops run eagerly / captured into a graph / replayed
    const std::pair<const char*, uint64_t> g[3] = {{"graph_eager", c->graphs.eager},
                                                   {"graph_captures", c->graphs.captures},
                                                   {"graph_replays", c->graphs.replays}};
    for (const auto& kv : g) {
        StageStat st{kv.first};
        st.items = (double)kv.second;
        all.push_back(st);
    }
    int n = (int)all.size();
    for (int i = 0; i < n && i < max; ++i) {
        memset(&out[i], 0, sizeof(out[i]));
        strncpy(out[i].name, all[i].name.c_str(), sizeof(out[i].name) - 1);
        out[i].ms = all[i].ms;
        out[i].launches = all[i].launches;
        out[i].alg_bytes = all[i].alg_bytes;
        out[i].items = all[i].items;
    }
    return n;
}

void curc_ctx_reset_stats(curc_ctx* c) {
    if (!c) return;
    c->stats.clear();
    c->graphs.eager = c->graphs.captures = c->graphs.replays = 0;
}

}  // extern "C"

extern "C" {

// Jobs run on up to CURC_BATCH_MAX_WORKERS streams of one device, each with
// its own scratch arena, fed largest job first from a shared counter (the
// reference instead starts one std::async task per block and serialises
// their GPU phases with a per-device try_lock, compress.cu:52-65,
// preprocess.cpp:205-276).
int curc_batch_run(curc_ctx* c, const curc_job* jobs, uint32_t n_jobs, uint32_t workers, curc_block** out,
                   int* status) {
    if (!c || (!jobs && n_jobs) || (!out && n_jobs)) return CURC_EINVAL;
    std::vector<uint64_t> w(n_jobs);
    for (uint32_t j = 0; j < n_jobs; ++j) w[j] = jobs[j].n_reads * (uint64_t)jobs[j].L;
    return run_on_workers(c, n_jobs, workers, w, out, status, [&](curc_ctx* wc, uint32_t j, curc_block** o) {
        return curc_block_run(wc, jobs[j].ascii, jobs[j].n_reads, jobs[j].L, CURC_STAGE_ALL, jobs[j].flags, o);
    });
}

// FASTQ input (preprocess.cpp:66-243): the record loop runs on the device.
int curc_fastq_block_run(curc_ctx* c, const curc_fastq_job* job, int stop_after, curc_block** out) {
    if (!c || !out) return CURC_EINVAL;
    *out = nullptr;
    return fq_block_run(c, *job, stop_after, S_FQ0, out, c->tune.fq_host);
}

int curc_fastq_batch_run(curc_ctx* c, const curc_fastq_job* jobs, uint32_t n_jobs, uint32_t workers,
                         curc_block** out, int* status) {
    if (!c || (!jobs && n_jobs) || (!out && n_jobs)) return CURC_EINVAL;
    std::vector<uint64_t> w(n_jobs);
    for (uint32_t j = 0; j < n_jobs; ++j) w[j] = jobs[j].bytes1 + (jobs[j].fq2 ? jobs[j].bytes2 : 0);
    // the host parse threads shared among the workers
    const uint32_t nw = std::max<uint32_t>(1, std::min<uint32_t>(workers ? workers : c->tune.workers, n_jobs));
    const uint32_t ht = c->tune.fq_host ? std::max<uint32_t>(1, c->tune.fq_host / nw) : 0;
    return run_on_workers(c, n_jobs, workers, w, out, status, [&](curc_ctx* wc, uint32_t j, curc_block** o) {
        return fq_block_run(wc, jobs[j], CURC_STAGE_ALL, S_FQ0, o, ht);
    });
}

// Blocks in order on the context's stream while a host thread copies the
// next block's bytes on the upload stream into the other of two device
// buffers (the reference parses the next chunks while block_compress runs,
// preprocess.cpp:91-243).  A copy into a buffer waits for the end of the
// parse of the block that used it before.
int curc_fastq_stream_run(curc_ctx* c, const curc_fastq_job* jobs, uint32_t n_jobs, curc_block_cb cb, void* user) {
    if (!c || (!jobs && n_jobs) || !cb) return CURC_EINVAL;
    for (uint32_t j = 0; j < n_jobs; ++j) {
        const int rc = fq_check_job(c, &jobs[j]);
        if (rc != CURC_OK) return rc;
        if (jobs[j].flags & CURC_RUN_INPUT_ON_DEVICE)
            return fail(c, CURC_EINVAL, "curc_fastq_stream_run copies host bytes (no CURC_RUN_INPUT_ON_DEVICE)");
    }
    if (n_jobs == 0) return CURC_OK;
    // host parse (tuning fq_host): the uploader parses block i into pinned
    // buffer i & 1 and copies only its sequences
    const uint32_t host_threads = c->tune.fq_host;
    uint8_t* buf[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> up(n_jobs, nullptr), used(n_jobs, nullptr);
    auto release = [&]() {
        for (uint32_t i = 0; i < n_jobs; ++i) {
            if (up[i]) hipEventDestroy(up[i]);
            if (used[i]) hipEventDestroy(used[i]);
        }
    };
    try {
        CURC_HIP(hipSetDevice(c->device));
        uint64_t mx = 0;
        for (uint32_t j = 0; j < n_jobs; ++j) mx = std::max(mx, fq_dev_bytes(jobs[j], host_threads));
        buf[0] = c->S.as<uint8_t>(S_FQ0, mx);
        buf[1] = c->S.as<uint8_t>(S_FQ1, mx);
        if (!c->ust) CURC_HIP(hipStreamCreateWithFlags(&c->ust, hipStreamNonBlocking));
        for (uint32_t i = 0; i < n_jobs; ++i) {
            CURC_HIP(hipEventCreateWithFlags(&up[i], hipEventDisableTiming));
            CURC_HIP(hipEventCreateWithFlags(&used[i], hipEventDisableTiming));
        }
    } catch (const std::exception& e) {
        release();
        return fail(c, CURC_EHIP, e.what());
    }
    std::mutex m;
    std::condition_variable cv;
    std::vector<int> uploaded(n_jobs, 0);  // 1: copies enqueued, -1: failed
    std::vector<char> consumed(n_jobs, 0); // the block's `used` event is recorded
    std::vector<FqInput> inputs(n_jobs);
    bool stop = false;
    std::string up_err;
    int up_code = CURC_EHIP;
    std::thread uploader([&]() {
        hipSetDevice(c->device);
        for (uint32_t i = 0; i < n_jobs; ++i) {
            if (i >= 2) {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return consumed[i - 2] || stop; });
                if (stop) return;
            }
            int ok = 1;
            try {
                if (i >= 2) {
                    // pinned buffer i & 1 is free once block i - 2's copies ended
                    if (host_threads) CURC_HIP(hipEventSynchronize(up[i - 2]));
                    CURC_HIP(hipStreamWaitEvent(c->ust, used[i - 2], 0));
                }
                inputs[i] = host_threads ? fq_stage_host(c, jobs[i], host_threads, (int)(i & 1), buf[i & 1], c->ust)
                                         : fq_upload(jobs[i], buf[i & 1], c->ust);
                CURC_HIP(hipEventRecord(up[i], c->ust));
            } catch (const std::exception& e) {
                ok = -1;
                up_err = e.what();
                if (dynamic_cast<const InputError*>(&e)) up_code = CURC_EINVAL;
            }
            {
                std::lock_guard<std::mutex> lk(m);
                uploaded[i] = ok;
            }
            cv.notify_all();
            if (ok < 0) return;
        }
    });
    int first = CURC_OK;
    for (uint32_t i = 0; i < n_jobs; ++i) {
        int state;
        {
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [&] { return uploaded[i] != 0; });
            state = uploaded[i];
        }
        curc_block* B = nullptr;
        int rc;
        if (state < 0) {
            rc = fail(c, up_code, up_err);
        } else {
            FqInput in = inputs[i];
            in.ready = up[i];
            in.done = used[i];
            rc = block_run(c, nullptr, 0, jobs[i].L, CURC_STAGE_ALL, fq_flags(jobs[i]), &in, &B);
        }
        {
            std::lock_guard<std::mutex> lk(m);
            consumed[i] = 1;
            if (rc != CURC_OK) stop = true;
        }
        cv.notify_all();
        cb(user, i, B, rc);
        if (rc != CURC_OK) {
            first = rc;
            break;
        }
    }
    {
        std::lock_guard<std::mutex> lk(m);
        stop = true;
    }
    cv.notify_all();
    uploader.join();
    stream_wait(c->ust);
    release();
    return first;
}

uint64_t curc_block_reads(const curc_block* B) { return B ? B->n_reads : 0; }

int curc_fastq_pack(curc_ctx* c, const char* d_fq, uint64_t bytes, uint32_t L, uint64_t max_reads,
                    uint64_t* d_packed, uint8_t* d_isN, uint64_t* d_seq_off, uint64_t* n_reads) {
    if (!c || !n_reads || (bytes && (!d_fq || !d_packed || !d_seq_off))) return CURC_EINVAL;
    *n_reads = 0;
    curc_params P;
    if (curc_make_params(L, &P) != CURC_OK) return fail(c, CURC_EINVAL, "unsupported read length");
    try {
        CURC_HIP(hipSetDevice(c->device));
        FqInput in;
        in.fq[0] = (const uint8_t*)d_fq;
        in.bytes[0] = bytes;
        uint32_t* cnt = c->S.as<uint32_t>(S_CNT, C_COUNT);
        *n_reads = fq_pack(c, in, L, P.W, max_reads, d_packed, d_isN, d_seq_off, cnt, nullptr, 0);
    } catch (const std::exception& e) {
        stream_wait(c->st);
        int code = CURC_EHIP;
        if (dynamic_cast<const InputError*>(&e)) code = CURC_EINVAL;
        else if (dynamic_cast<const std::length_error*>(&e)) code = CURC_ELIMIT;
        return fail(c, code, e.what());
    }
    return CURC_OK;
}

// curc_fastq_pack without allocation: counters and tile counts live in the
// fixed device array `scratch` (through its per-device address), so no
// device or host memory is allocated.  Calls that overlap on one GPU share
// that array and race.
int curc_fastq_pack_static(curc_ctx* c, const char* d_fq, uint64_t bytes, uint32_t L, uint64_t max_reads,
                           uint64_t* d_packed, uint8_t* d_isN, uint64_t* d_seq_off, uint64_t* n_reads) {
    if (!c || !n_reads || (bytes && (!d_fq || !d_packed || !d_seq_off))) return CURC_EINVAL;
    *n_reads = 0;
    curc_params P;
    if (curc_make_params(L, &P) != CURC_OK) return fail(c, CURC_EINVAL, "unsupported read length");
    if (bytes > CURC_FASTQ_STATIC_MAX_BYTES)
        return fail(c, CURC_ELIMIT, "FASTQ bytes beyond the fixed scratch maximum (CURC_FASTQ_STATIC_MAX_BYTES)");
    uint64_t* base = static_scratch_init(c->device);  // resolved when the context was created
    if (!base) return fail(c, CURC_EHIP, "fixed scratch not resolved on this device");
    try {
        CURC_HIP(hipSetDevice(c->device));
        FqInput in;
        in.fq[0] = (const uint8_t*)d_fq;
        in.bytes[0] = bytes;
        uint32_t* cnt = reinterpret_cast<uint32_t*>(base);
        *n_reads = fq_pack(c, in, L, P.W, max_reads, d_packed, d_isN, d_seq_off, cnt, base + STATIC_CNT_WORDS,
                           STATIC_SCRATCH_WORDS - STATIC_CNT_WORDS);
    } catch (const std::exception& e) {
        stream_wait(c->st);
        int code = CURC_EHIP;
        if (dynamic_cast<const InputError*>(&e)) code = CURC_EINVAL;
        else if (dynamic_cast<const std::length_error*>(&e)) code = CURC_ELIMIT;
        return fail(c, code, e.what());
    }
    return CURC_OK;
}

int curc_decode_block(curc_ctx* c, const curc_decode_job* j, char* out1, char* out2) {
    if (!c || !j || (j->n_reads && !out1)) return CURC_EINVAL;
    if (j->L == 0 || j->L > 512) return fail(c, CURC_EINVAL, "unsupported read length");
    if ((j->flags & CURC_RUN_PAIRED) && j->n_reads && !out2) return fail(c, CURC_EINVAL, "paired output needs out2");
    if (j->flags & CURC_RUN_INPUT_ON_DEVICE) return fail(c, CURC_EINVAL, "decode inputs are host arrays");
    if (j->max_mm > 255 || (j->max_mm && (!j->mis_off || !j->mis_off_len)))
        return fail(c, CURC_EINVAL, "mismatch-offset streams missing");
    try {
        CURC_HIP(hipSetDevice(c->device));
        decode_run(c, *j, out1, out2);
        c->harvest();
    } catch (const std::exception& e) {
        stream_wait(c->st);
        c->pending.clear();
        int code = CURC_EHIP;
        if (dynamic_cast<const InputError*>(&e)) code = CURC_EINVAL;
        else if (std::string(e.what()).find("out of memory") != std::string::npos) code = CURC_ENOMEM;
        return fail(c, code, e.what());
    }
    return CURC_OK;
}

}  // extern "C"

// ------------------------------------------------------------ scorer net ---
struct curc_mlp {
    curc_ctx* c = nullptr;
    MlpDev d{};
    uint64_t t = 0;  // Adam steps taken
    // training intermediates, grown on demand
    uint64_t cap = 0;
    uint16_t *xt = nullptr, *h1t = nullptr, *dz2t = nullptr, *dz1t = nullptr;
    float *h2 = nullptr, *ds = nullptr, *logit = nullptr, *lsum = nullptr;
    void grow(uint64_t n) {
        if (n <= cap) return;
        free_tmp();
        cap = n + n / 8 + 64;
        CURC_HIP(hipMalloc(&xt, cap * MLP_IN * 2));
        CURC_HIP(hipMalloc(&h1t, cap * MLP_H * 2));
        CURC_HIP(hipMalloc(&dz2t, cap * MLP_H * 2));
        CURC_HIP(hipMalloc(&dz1t, cap * MLP_H * 2));
        CURC_HIP(hipMalloc(&h2, cap * MLP_H * 4));
        CURC_HIP(hipMalloc(&ds, cap * 4));
        CURC_HIP(hipMalloc(&logit, cap * 4));
    }
    void free_tmp() {
        for (void* q : {(void*)xt, (void*)h1t, (void*)dz2t, (void*)dz1t, (void*)h2, (void*)ds, (void*)logit})
            if (q) hipFree(q);
        xt = h1t = dz2t = dz1t = nullptr;
        h2 = ds = logit = nullptr;
        cap = 0;
    }
};

namespace {

// Forward, loss gradient and all parameter gradients of the mean
// cross-entropy over n samples, accumulated into d.g.
void mlp_backprop(curc_mlp* m, const float* x, const float* y, uint64_t n) {
    hipStream_t st = m->c->st;
    m->grow(n);
    const uint32_t g = mlp_grid(n);
    hipLaunchKernelGGL(k_mlp_train_fwd, dim3(g), dim3(256), 0, st, m->d, x, y, n, 1.0f / (float)n, m->xt, m->h1t,
                       m->h2, m->ds, m->logit);
    hipLaunchKernelGGL(k_mlp_train_bwd, dim3(g), dim3(256), 0, st, m->d, n, (const float*)m->h2,
                       (const float*)m->ds, (const uint16_t*)m->h1t, m->dz2t, m->dz1t);
    const uint32_t parts = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(64, n / 512));
    hipLaunchKernelGGL(k_mlp_train_wgrad, dim3(20, parts), dim3(64), 0, st, m->d, n, (const uint16_t*)m->xt,
                       (const uint16_t*)m->h1t, (const uint16_t*)m->dz2t, (const uint16_t*)m->dz1t, parts);
    hipLaunchKernelGGL(k_mlp_train_b1, dim3(MLP_H), dim3(256), 0, st, m->d, n, (const uint16_t*)m->dz1t);
}

}  // namespace

extern "C" {

curc_mlp* curc_mlp_create(curc_ctx* c) {
    if (!c) return nullptr;
    curc_mlp* m = new curc_mlp();
    m->c = c;
    try {
        CURC_HIP(hipSetDevice(c->device));
        for (float** q : {&m->d.p, &m->d.m, &m->d.v, &m->d.g}) CURC_HIP(hipMalloc(q, MLP_NPARAMS * 4));
        CURC_HIP(hipMalloc(&m->d.w1t, MLP_IN * MLP_H * 2));
        CURC_HIP(hipMalloc(&m->d.w2t, MLP_H * MLP_H * 2));
        CURC_HIP(hipMalloc(&m->d.w2, MLP_H * MLP_H * 2));
        CURC_HIP(hipMalloc(&m->lsum, 4));
        for (float* q : {m->d.p, m->d.m, m->d.v, m->d.g})
            hipLaunchKernelGGL(k_fill<float>, dim3(32), dim3(256), 0, c->st, q, (uint64_t)MLP_NPARAMS, 0.f);
        hipLaunchKernelGGL(k_mlp_pack, dim3(64), dim3(256), 0, c->st, m->d);
        CURC_HIP(stream_wait(c->st));
    } catch (const std::exception& e) {
        c->err = e.what();
        curc_mlp_destroy(m);
        return nullptr;
    }
    return m;
}

void curc_mlp_destroy(curc_mlp* m) {
    if (!m) return;
    // the context may already be gone: hipFree waits for pending work itself
    m->free_tmp();
    for (void* q : {(void*)m->d.p, (void*)m->d.m, (void*)m->d.v, (void*)m->d.g, (void*)m->d.w1t, (void*)m->d.w2t,
                    (void*)m->d.w2, (void*)m->lsum})
        if (q) hipFree(q);
    delete m;
}

int curc_mlp_set_params(curc_mlp* m, const float* params, uint32_t n) {
    if (!m || !params || n != MLP_NPARAMS) return CURC_EINVAL;
    try {
        hipStream_t st = m->c->st;
        CURC_HIP(hipMemcpyAsync(m->d.p, params, MLP_NPARAMS * 4, hipMemcpyHostToDevice, st));
        for (float* q : {m->d.m, m->d.v, m->d.g})
            hipLaunchKernelGGL(k_fill<float>, dim3(32), dim3(256), 0, st, q, (uint64_t)MLP_NPARAMS, 0.f);
        hipLaunchKernelGGL(k_mlp_pack, dim3(64), dim3(256), 0, st, m->d);
        CURC_HIP(stream_wait(st));
        m->t = 0;
    } catch (const std::exception& e) {
        return fail(m->c, CURC_EHIP, e.what());
    }
    return CURC_OK;
}

int curc_mlp_get_params(curc_mlp* m, float* params, uint32_t n) {
    if (!m || !params || n != MLP_NPARAMS) return CURC_EINVAL;
    try {
        CURC_HIP(hipMemcpyAsync(params, m->d.p, MLP_NPARAMS * 4, hipMemcpyDeviceToHost, m->c->st));
        CURC_HIP(stream_wait(m->c->st));
    } catch (const std::exception& e) {
        return fail(m->c, CURC_EHIP, e.what());
    }
    return CURC_OK;
}

int curc_mlp_forward(curc_mlp* m, const float* d_x, uint64_t n, float* d_logit) {
    if (!m || (n && (!d_x || !d_logit))) return CURC_EINVAL;
    if (n == 0) return CURC_OK;
    try {
        hipLaunchKernelGGL(k_mlp_infer, dim3(mlp_infer_grid(n)), dim3(64 * MLP_INF_WAVES), 0, m->c->st, m->d, d_x, n,
                           (const uint32_t*)nullptr, d_logit);
        CURC_HIP(hipGetLastError());
        CURC_HIP(stream_wait(m->c->st));
    } catch (const std::exception& e) {
        return fail(m->c, CURC_EHIP, e.what());
    }
    return CURC_OK;
}

int curc_mlp_gradients(curc_mlp* m, const float* d_x, const float* d_y, uint64_t n, float* grads) {
    if (!m || !d_x || !d_y || !grads || n == 0) return CURC_EINVAL;
    try {
        hipStream_t st = m->c->st;
        hipLaunchKernelGGL(k_fill<float>, dim3(32), dim3(256), 0, st, m->d.g, (uint64_t)MLP_NPARAMS, 0.f);
        mlp_backprop(m, d_x, d_y, n);
        CURC_HIP(hipMemcpyAsync(grads, m->d.g, MLP_NPARAMS * 4, hipMemcpyDeviceToHost, st));
        hipLaunchKernelGGL(k_fill<float>, dim3(32), dim3(256), 0, st, m->d.g, (uint64_t)MLP_NPARAMS, 0.f);
        CURC_HIP(hipGetLastError());
        CURC_HIP(stream_wait(st));
    } catch (const std::exception& e) {
        return fail(m->c, CURC_EHIP, e.what());
    }
    return CURC_OK;
}

int curc_mlp_train_step(curc_mlp* m, const float* d_x, const float* d_y, uint64_t n, float lr, float* loss) {
    if (!m || !d_x || !d_y || n == 0) return CURC_EINVAL;
    try {
        hipStream_t st = m->c->st;
        mlp_backprop(m, d_x, d_y, n);
        if (loss) {
            hipLaunchKernelGGL(k_fill<float>, dim3(1), dim3(64), 0, st, m->lsum, (uint64_t)1, 0.f);
            hipLaunchKernelGGL(k_bce_sum, dim3((uint32_t)std::min<uint64_t>(1024, (n + 255) / 256)), dim3(256), 0,
                               st, (const float*)m->logit, d_y, n, m->lsum);
        }
        m->t += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)m->t)));
        const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)m->t)));
        hipLaunchKernelGGL(k_mlp_adam, dim3(32), dim3(256), 0, st, m->d, lr, b1, b2, 1e-8f, c1, c2);
        hipLaunchKernelGGL(k_mlp_pack, dim3(64), dim3(256), 0, st, m->d);
        CURC_HIP(hipGetLastError());
        if (loss) {
            float sum = 0;
            CURC_HIP(hipMemcpyAsync(&sum, m->lsum, 4, hipMemcpyDeviceToHost, st));
            CURC_HIP(stream_wait(st));
            *loss = sum / (float)n;
        }
    } catch (const std::exception& e) {
        return fail(m->c, CURC_EHIP, e.what());
    }
    return CURC_OK;
}

}  // extern "C"

static_assert(CURC_MLP_NPARAMS == curc::MLP_NPARAMS, "scorer parameter count");

extern "C" {

int curc_ctx_set_scorer(curc_ctx* c, curc_mlp* m, float threshold_mapped, float threshold_unmapped) {
    if (!c) return CURC_EINVAL;
    c->has_scorer = m != nullptr;
    if (m) c->scorer_net = m->d;
    c->scorer_threshold = threshold_mapped;
    c->scorer_threshold_unmapped = threshold_unmapped;
    return CURC_OK;
}

int curc_ctx_set_trace(curc_ctx* c, uint32_t every, uint64_t max_records) {
    if (!c || every == 0 || max_records == 0) return CURC_EINVAL;
    c->trace_every = every;
    c->trace_cap = max_records;
    return CURC_OK;
}

int curc_ctx_get_tuning(curc_ctx* c, curc_tuning* t) {
    if (!c || !t) return CURC_EINVAL;
    t->ov_grid = c->tune.ov_grid;
    t->fm_grid = c->tune.fm_grid;
    t->map_grid = c->tune.map_grid;
    t->radix_items = c->tune.radix_items;
    t->workers = c->tune.workers;
    t->filter_bits = c->tune.filter_bits;
    t->fm_fused = c->tune.fm_fused;
    t->ov_queue = c->tune.ov_queue;
    t->map_both = c->tune.map_both;
    t->fm_static = c->tune.fm_static;
    t->fq_host = c->tune.fq_host;
    return CURC_OK;
}

int curc_ctx_set_tuning(curc_ctx* c, const curc_tuning* t) {
    if (!c || !t) return CURC_EINVAL;
    if (!t->ov_grid || !t->fm_grid || !t->map_grid || !t->workers || t->workers > CURC_BATCH_MAX_WORKERS ||
        (t->radix_items != 8 && t->radix_items != 16 && t->radix_items != 32) || t->filter_bits < 8 ||
        t->filter_bits > 128 || (t->filter_bits & (t->filter_bits - 1)) || t->ov_queue > 3 || t->map_both > 2 ||
        t->fm_static > 1 || t->fq_host > CURC_FQ_HOST_MAX)
        return fail(c, CURC_EINVAL, "tuning value out of range");
    c->tune.ov_grid = t->ov_grid;
    c->tune.fm_grid = t->fm_grid;
    c->tune.map_grid = t->map_grid;
    c->tune.radix_items = t->radix_items;
    c->tune.workers = t->workers;
    c->tune.filter_bits = t->filter_bits;
    c->tune.fm_fused = t->fm_fused;
    c->tune.ov_queue = t->ov_queue;
    c->tune.map_both = t->map_both;
    c->tune.fm_static = t->fm_static;
    c->tune.fq_host = t->fq_host;
    return CURC_OK;
}

}  // extern "C"

extern "C" {

// RefMatcher on its own (compress.cu:630-1001): init on src (the HQ
// pseudogenome, A/C/G/T) and one match of dest; results in the block's
// CURC_F_RM_* fields of part 0 (matches and found with KEEP_STAGES).
int curc_ref_match(curc_ctx* c, const char* src, uint64_t src_len, const char* dest, uint64_t dest_len,
                   int dest_contain_n, int dest_is_ref, uint32_t flags, curc_block** out) {
    if (!c || !out || (src_len && !src) || (dest_len && !dest)) return CURC_EINVAL;
    *out = nullptr;
    curc_block* B = new curc_block();
    try {
        CURC_HIP(hipSetDevice(c->device));
        radix_items() = c->tune.radix_items;
        Scratch& S = c->S;
        hipStream_t st = c->st;
        uint32_t* cnt = S.as<uint32_t>(S_CNT, C_COUNT);
        const uint64_t sw = (src_len + 31) / 32 + 1;
        uint8_t* dsrc = S.as<uint8_t>(S_ASCII, src_len + 1);
        uint64_t* src2 = S.as<uint64_t>(S_PG_HQ, sw);
        h2d(c, dsrc, (const uint8_t*)src, src_len);
        hipLaunchKernelGGL(k_rm_pack_text, dim3(grid_for(sw)), dim3(256), 0, st, (const uint8_t*)dsrc, src_len, src2,
                           sw);
        const uint64_t dw = (dest_len + 31) / 32 + 1;
        uint8_t* ddest = S.as<uint8_t>(S_NPG, dest_len + 16);
        h2d(c, ddest, (const uint8_t*)dest, dest_len);
        uint64_t* dest2 = nullptr;
        if (!dest_contain_n) {
            dest2 = S.as<uint64_t>(S_PG_LQ, dw);
            hipLaunchKernelGGL(k_rm_pack_text, dim3(grid_for(dw)), dim3(256), 0, st, (const uint8_t*)ddest, dest_len,
                               dest2, dw);
        }
        Region r(c, "refmatch");
        const RmIndex ix = rm_index(c, src2, src_len);
        rm_match(c, ix, src2, src_len, dest2, dest_contain_n ? ddest : nullptr, dest_len, dest_is_ref != 0,
                 (flags & CURC_RUN_KEEP_STAGES) != 0, cnt, B->rm[0]);
        CURC_HIP(hipGetLastError());
        CURC_HIP(stream_wait(st));
        c->harvest();
    } catch (const std::length_error& e) {
        // copies into B's arrays may be in flight: join them before freeing
        stream_wait(c->st);
        if (c->cst) stream_wait(c->cst);
        delete B;
        c->pending.clear();
        return fail(c, CURC_ELIMIT, e.what());
    } catch (const std::exception& e) {
        stream_wait(c->st);
        if (c->cst) stream_wait(c->cst);
        delete B;
        c->pending.clear();
        return fail(c, CURC_EHIP, e.what());
    }
    *out = B;
    return CURC_OK;
}

}  // extern "C"
