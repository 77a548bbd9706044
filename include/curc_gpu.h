/*
 * libcurc_gpu: MI355X (gfx950) implementation of CURC's pseudogenome hot path
 * behind a C ABI.
 *
 * The reference's GPU/host boundary is src/compress.cuh (C++ linkage):
 *   block_compress(...)             compress.cuh:48-54
 *   encode(...)                     compress.cuh:56-57
 *   get_device_count_and_init(...)  compress.cuh:59
 *   gpu_malloc_host / gpu_malloc / gpu_free_host / gpu_free /
 *   gpu_mem_H2D / gpu_mem_D2H       compress.cuh:61-71
 * Each declaration below names the reference entry it replaces.
 *
 * Conventions: every function returns an int status (CURC_OK = 0) and never
 * exits the process; the caller owns every buffer it passes in, and the
 * library never frees caller memory; results land in caller buffers or in a
 * curc_block handle released with curc_block_free.  One context per host
 * thread; contexts are independent.
 */
#ifndef CURC_GPU_H
#define CURC_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CURC_OK 0
#define CURC_EINVAL -1 /* bad argument or unsupported input (e.g. L > 512) */
#define CURC_EHIP -2   /* HIP runtime error; text in curc_ctx_last_error */
#define CURC_ENOMEM -3 /* device or host allocation failed */
#define CURC_ELIMIT -4 /* input exceeds a documented limit */

/* Derived per-read-length parameters (preprocess.cpp:296-313, Param.hpp:47-76). */
typedef struct curc_params {
    uint32_t L;       /* read length, 1..512 */
    uint32_t W;       /* 64-bit words per packed read, ceil(L/32) */
    uint32_t max_off; /* floor(65 L / 100) */
    uint32_t k;       /* overlap k-mer: 32 if L >= 89 else floor(36 L / 100) */
    uint32_t kip;     /* k-mer index position, L - k - max_off */
    uint32_t max_mm;  /* mapping mismatch limit, floor(L / 5) */
} curc_params;

int curc_make_params(uint32_t L, curc_params *p);

/* 16-byte record, layout-compatible with AlignmentRecord (compress.cu:68-77):
 * word = is_contain_N | strand_id << 1 | pos << 2. */
typedef struct curc_alignment {
    uint32_t read_id;
    uint32_t pad;
    uint64_t word;
} curc_alignment;

typedef struct curc_ctx curc_ctx;

/* Replaces get_device_count_and_init (compress.cu:2183-2193). */
int curc_device_count(int *count);
/* One context per device and host thread; holds a stream and device scratch. */
curc_ctx *curc_ctx_create(int device);
void curc_ctx_destroy(curc_ctx *ctx);
const char *curc_ctx_last_error(curc_ctx *ctx);
/* Drops cached device scratch (it is otherwise kept for the next block) and
 * returns the pinned host chunks that no live block holds.  Large block
 * results live in pinned host memory that is recycled across blocks. */
int curc_ctx_trim(curc_ctx *ctx);

/* Memory helpers; replace gpu_malloc_host / gpu_malloc / gpu_free_host /
 * gpu_free / gpu_mem_H2D / gpu_mem_D2H (compress.cu:2195-2220). */
int curc_host_alloc(curc_ctx *ctx, void **ptr, uint64_t bytes);
int curc_host_free(curc_ctx *ctx, void *ptr);
int curc_dev_alloc(curc_ctx *ctx, void **ptr, uint64_t bytes);
int curc_dev_free(curc_ctx *ctx, void *ptr);
int curc_memcpy_h2d(curc_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int curc_memcpy_d2h(curc_ctx *ctx, void *dst, const void *src, uint64_t bytes);
/* No counterpart in the reference.  Copies `bytes` (a multiple of 16, both
 * pointers 16-byte aligned, device memory) `reps` times with a plain
 * 16-byte-word copy kernel and returns the mean time per copy in *ms: the
 * measured bandwidth ceiling beside the vendor peak. */
int curc_stream_copy(curc_ctx *ctx, void *dst, const void *src, uint64_t bytes, int reps, float *ms);
int curc_sync(curc_ctx *ctx);
/* Waits for all work on the context's device (hipDeviceSynchronize). */
int curc_device_sync(curc_ctx *ctx);

/* 2-bit pack of n ASCII reads of length L (all device pointers); replaces
 * encode (compress.cu:2151-2180).  d_packed holds n*ceil(L/32) words; d_isN
 * (optional) receives 1 for reads containing 'N' (preprocess.cpp:151). */
int curc_pack(curc_ctx *ctx, const char *d_ascii, uint64_t n, uint32_t L, uint64_t *d_packed,
              uint8_t *d_isN);

/* Stages of one block (compress.cu:1349-1728).  curc_block_run stops after
 * the requested stage. */
enum {
    CURC_STAGE_DEDUP = 1,     /* a4: duplicate links */
    CURC_STAGE_OVERLAP = 2,   /* a5-a6: prefix index + greedy overlap */
    CURC_STAGE_PRUNE = 3,     /* a7: prune + HQ/unmatched partition */
    CURC_STAGE_FULLMATCH = 4, /* a8: exact all-offset matching of HQ ends */
    CURC_STAGE_ASSEMBLY = 5,  /* a9: HQ pseudogenome */
    CURC_STAGE_MAP = 6,       /* a10: mapping onto the HQ pseudogenome */
    CURC_STAGE_LQ = 7,        /* a11: LQ pseudogenome of unmapped reads */
    CURC_STAGE_NLQ = 8,       /* a12: N-read LQ pseudogenome of unmapped N-reads */
    CURC_STAGE_ALL = 8
};

/* Result fields (numbering shared with the oracle's field ids). */
enum {
    CURC_F_PACKED = 0,         /* u64[n_reads * W] */
    CURC_F_NONN_IDS = 1,       /* u32 */
    CURC_F_N_IDS = 2,          /* u32 */
    CURC_F_SORTED_IDS = 3,     /* u32 non-N ids in dedup order */
    CURC_F_NEXT_DEDUP = 4,     /* u32[n_reads] */
    CURC_F_NEXT_OVERLAP = 5,   /* u32[n_reads] */
    CURC_F_PREV_OVERLAP = 6,   /* u32[n_reads] */
    CURC_F_OFFSET_OVERLAP = 7, /* u16[n_reads] */
    CURC_F_NEXT = 8,           /* u32[n_reads] final */
    CURC_F_PREV = 9,
    CURC_F_OFFSET = 10,        /* u16 */
    CURC_F_HQ_IDS = 11,        /* u32 */
    CURC_F_UNMATCHED_IDS = 12, /* u32 */
    CURC_F_HQ_PG = 13,         /* char, HQ pseudogenome text */
    CURC_F_HQ_REC = 14,        /* curc_alignment */
    CURC_F_MAP_REC = 15,       /* curc_alignment */
    CURC_F_MAP_MIS = 16,       /* u8 mismatch count per map record */
    CURC_F_UNMAPPED = 17,      /* u32 global ids of unmapped non-N reads */
    CURC_F_UNMAPPED_N = 18,    /* u32 N indices of unmapped N-reads */
    CURC_F_LQ_PG = 19,         /* char */
    CURC_F_LQ_REC = 20,        /* curc_alignment, global ids */
    CURC_F_NLQ_PG = 21,        /* char, N-read LQ pseudogenome text */
    CURC_F_NLQ_REC = 22,       /* curc_alignment, N indices, is_contain_N set */
    CURC_F_PREV_DEDUP = 23,
    CURC_F_NEXT_PRUNE = 24,
    CURC_F_PREV_PRUNE = 25,
    CURC_F_HQ_PG2 = 26,        /* u64 2-bit HQ pseudogenome, ceil(len/32)+1 words */
    CURC_F_LQ_NEXT = 27,       /* u32, local ids of the LQ graph */
    CURC_F_LQ_PREV = 28,
    CURC_F_LQ_OFFSET = 29,     /* u16 */
    CURC_F_N_ASCII = 30,       /* char[n_N * L] N-reads in N-index order */
    /* pre-entropy streams (CURC_RUN_STREAMS; compress.cu:1730-2044) */
    CURC_F_S_STRAND = 40,      /* char '0'/'1' per sorted HQ+map record */
    CURC_F_S_READ_OFF = 41,    /* u16 position delta per sorted record (order not preserved) */
    CURC_F_S_MIS_COUNT = 42,   /* u8 mismatch count per sorted record */
    CURC_F_S_MIS_BASE = 43,    /* u8 recoded mismatch symbol per mismatch */
    CURC_F_S_MIS_TABLE = 44,   /* 16 bytes: decoder rows for contexts A, T, C, G */
    CURC_F_S_LQ_OFF = 45,      /* u16 L - position delta per LQ record */
    CURC_F_S_NLQ_OFF = 46,     /* u16 L - position delta per N-LQ record */
    CURC_F_S_PE_ORDER = 47,    /* u32 global ids in emission order (paired, order not preserved) */
    CURC_F_S_ID_POS = 48,      /* u64 joined-pseudogenome position per read (order preserved) */
    CURC_F_S_SORTED_REC = 49,  /* curc_alignment, HQ+map records after the record sort */
    /* traced mapping (CURC_RUN_TRACE): one record per sampled read and checkpoint */
    CURC_F_TRACE_X = 60,       /* float32 [n][32] scorer features (count: n records) */
    CURC_F_TRACE_Y = 61,       /* float32 [n]: 1 if the rest of the strand pass lowered the best count */
    CURC_F_S_MIS_OFF = 64,     /* + m (1..max_mm): u16 mismatch-offset stream for m mismatches */
    /* pseudogenome self-match (CURC_RUN_REFMATCH; RefMatcher, compress.cu:597-1001, called at
     * :2053-2069).  Part p (0 = HQ text against itself, 1 = LQ text, 2 = N-read LQ text, each
     * against the HQ text) adds 5p to: */
    CURC_F_RM_BASE = 1000,
    CURC_F_RM_TEXT = 1000,     /* char: text with each reference replaced by '%' */
    CURC_F_RM_MAP_OFF = 1001,  /* bytes: u32 source positions (u64 if the HQ text exceeds 2^32-1) */
    CURC_F_RM_MAP_LEN = 1002,  /* bytes: byte-frugal 50, then byte-frugal (length - 50) per reference */
    CURC_F_RM_MATCHES = 1003,  /* u64 (src, len, dest) triples after sort + unique (KEEP_STAGES) */
    CURC_F_RM_FOUND = 1004     /* u64 triples in collection order, before the capacity cut (KEEP_STAGES) */
};

/* Flags for curc_block_run. */
#define CURC_RUN_INPUT_ON_DEVICE 1u /* `ascii` is a device pointer */
#define CURC_RUN_KEEP_STAGES 2u     /* also return the per-stage link snapshots */
#define CURC_RUN_STREAMS 4u         /* also build the block's pre-entropy streams */
#define CURC_RUN_PRESERVE_ORDER 8u  /* streams for --preserve_order */
#define CURC_RUN_PAIRED 16u         /* streams for paired-end input (ids 2k / 2k+1) */
#define CURC_RUN_REFMATCH 32u       /* also self-match the pseudogenomes (CURC_F_RM_*) */
#define CURC_RUN_TRACE 64u          /* record scorer features and labels while mapping (CURC_F_TRACE_*) */
#define CURC_RUN_LEARNED 128u       /* map with the context's scorer (curc_ctx_set_scorer) */

typedef struct curc_block curc_block;

/* Pseudogenome part of block_compress (compress.cu:1349-1728): N-split and
 * pack, dedup, prefix index, greedy overlap, prune/partition, exact
 * all-offset matching, assembly, mapping, the LQ pseudogenome and the N-read
 * LQ pseudogenome (which the reference runs on the CPU, full_match_cpu), for
 * n_reads ASCII reads of length L (block-global ids = input order). */
int curc_block_run(curc_ctx *ctx, const char *ascii, uint64_t n_reads, uint32_t L,
                   int stop_after, uint32_t flags, curc_block **out);
uint64_t curc_block_get(curc_block *b, int field, const void **ptr);
void curc_block_free(curc_block *b);

/* Per-stage timing (HIP events on the context's stream) and algorithmic
 * byte counts accumulated since the last reset. */
typedef struct curc_stage_stat {
    char name[32];
    double ms;         /* summed event time */
    uint64_t launches; /* timed regions recorded */
    double alg_bytes;  /* algorithmic bytes moved by the stage */
    double items;      /* work items (probes, candidates, ...) */
} curc_stage_stat;

int curc_ctx_set_timing(curc_ctx *ctx, int on);
int curc_ctx_stats(curc_ctx *ctx, curc_stage_stat *out, int max_stats);
void curc_ctx_reset_stats(curc_ctx *ctx);

/* Seeded synthetic workloads (no reference counterpart; test/bench input). */
typedef struct curc_synth_cfg {
    uint64_t seed;
    uint64_t genome_len;
    double gc;             /* P(G)+P(C) */
    double repeat_frac;    /* fraction of the genome covered by repeat copies */
    uint32_t rep_min, rep_max;
    double div_min, div_max;
    uint32_t rep_families;
    uint64_t n_reads;      /* total reads (pairs interleaved when paired) */
    uint32_t L;
    int32_t paired;
    double frag_mean, frag_sd;
    double rc_prob;
    double sub_first, sub_last; /* substitution rate at base 1 and base L, linear between */
    double n_rate;
} curc_synth_cfg;

void curc_synth_default(curc_synth_cfg *cfg);
int curc_synth_genome(const curc_synth_cfg *cfg, char *genome);
int curc_synth_reads(const curc_synth_cfg *cfg, const char *genome, char *reads);
/* FASTQ records of n reads (read i at reads + i * stride): header
 * "@<prefix><first_id + i, 10 digits>" (with "/<mate>" when mate > 0), the
 * sequence, "+", L quality bytes 'I'; every record is
 * curc_synth_fastq_record_len bytes. */
uint64_t curc_synth_fastq_record_len(uint32_t L, const char *prefix, int mate);
int curc_synth_fastq(const char *reads, uint64_t n, uint32_t L, uint64_t stride, uint64_t first_id,
                     const char *prefix, int mate, char *out);

/* RefMatcher on its own (compress.cu:630-1001): RefMatcher::init on src
 * (the HQ pseudogenome text) and one RefMatcher::match of dest (a text with
 * N when dest_contain_n, the character path; dest_is_ref for the self match
 * of src).  Results in the block's CURC_F_RM_* fields of part 0 (found and
 * matches with CURC_RUN_KEEP_STAGES in flags). */
int curc_ref_match(curc_ctx *ctx, const char *src, uint64_t src_len, const char *dest, uint64_t dest_len,
                   int dest_contain_n, int dest_is_ref, uint32_t flags, curc_block **out);

/* ---- batching ------------------------------------------------------------
 * Many independent blocks or inputs (uneven read counts and lengths) in one
 * call.  Replaces the per-block std::async fan-out of preprocess.cpp:205-276
 * whose GPU phases serialise on a per-device try_lock (compress.cu:52-65):
 * here the jobs run concurrently on up to CURC_BATCH_MAX_WORKERS streams of
 * the context's device, each with its own scratch arena, largest job first.
 * out[i] receives job i's block (NULL if it failed; status[i], when given,
 * holds its code).  workers = 0 picks the maximum.  Returns the first
 * failing job's code, or CURC_OK. */
#define CURC_BATCH_MAX_WORKERS 4u
typedef struct curc_job {
    const char *ascii; /* n_reads * L ASCII bytes (device pointer with CURC_RUN_INPUT_ON_DEVICE) */
    uint64_t n_reads;
    uint32_t L;
    uint32_t flags; /* CURC_RUN_* */
} curc_job;
int curc_batch_run(curc_ctx *ctx, const curc_job *jobs, uint32_t n_jobs, uint32_t workers, curc_block **out,
                   int *status);

/* ---- FASTQ input -----------------------------------------------------------
 * The record loop of the reference's host parser (preprocess.cpp:147-199,
 * with the read length of compress(), :281-292) runs on the device: a
 * block's FASTQ bytes are copied to HBM as they are and parsed there (four
 * lines per record, the sequence line kept, 'N' reads split off, 2-bit
 * packed).  A block is a run of whole records; where blocks end is the
 * caller's chunking (preprocess.cpp:91-138, 200-243; curc_amd/fastq.py).
 * Malformed input (a sequence line of another length than L: "read length
 * isn't fix", :154; a block not starting with '@' or ending mid-record;
 * paired files with different record counts) fails with CURC_EINVAL.  Block
 * ids follow record order; a paired block (fq2 given) interleaves file 1's
 * record k as read 2k and file 2's as 2k+1 (:156-192), and its streams are
 * built as paired. */
typedef struct curc_fastq_job {
    const char *fq1;  /* file 1's bytes of the block (host memory; device with CURC_RUN_INPUT_ON_DEVICE,
                         then 16-byte aligned) */
    uint64_t bytes1;
    const char *fq2;  /* file 2's bytes of a paired block, else NULL */
    uint64_t bytes2;
    uint32_t L;       /* read length of the input */
    uint32_t flags;   /* CURC_RUN_* */
} curc_fastq_job;

/* One block from FASTQ bytes; replaces the per-chunk parse, encode and
 * block_compress calls of process() (preprocess.cpp:147-213) for one block.
 * Host bytes are copied to the device first (pinned memory copies fastest). */
int curc_fastq_block_run(curc_ctx *ctx, const curc_fastq_job *job, int stop_after, curc_block **out);
/* Blocks from FASTQ bytes on up to CURC_BATCH_MAX_WORKERS streams, as
 * curc_batch_run (each worker copies its block's bytes, then runs it). */
int curc_fastq_batch_run(curc_ctx *ctx, const curc_fastq_job *jobs, uint32_t n_jobs, uint32_t workers,
                         curc_block **out, int *status);
/* Blocks from FASTQ bytes in host memory, in order on the context's stream,
 * while the next block's bytes are copied to the device beside it (the
 * reference parses the next chunks while block_compress runs,
 * preprocess.cpp:91-243).  Each finished block goes to cb on the calling
 * thread, in job order, with its status; cb owns it (curc_block_free).
 * Stops at the first failing block (its cb gets a NULL block) and returns
 * its code. */
typedef void (*curc_block_cb)(void *user, uint32_t job, curc_block *block, int status);
int curc_fastq_stream_run(curc_ctx *ctx, const curc_fastq_job *jobs, uint32_t n_jobs, curc_block_cb cb,
                          void *user);
/* Reads in a block (the parsed record count for FASTQ input). */
uint64_t curc_block_reads(const curc_block *b);
/* The record loop of preprocess.cpp:147-199 on the host, for one file's
 * bytes (no device needed): the sequence lines, L bytes each, in record
 * order, into out (out_bytes >= curc_fastq_host_cap(bytes, L, threads)),
 * using `threads` threads.  *format_error gets 0, or the first rule the
 * bytes break, in the device parser's order: 1 they do not start with '@',
 * 2 they end inside a record, 3 a sequence line does not hold L bytes
 * (:154); then CURC_EINVAL is returned.  curc_fastq_stream_run and
 * curc_fastq_block_run parse host bytes this way when the tuning field
 * fq_host is not 0, and copy only the sequences to the device. */
uint64_t curc_fastq_host_cap(uint64_t bytes, uint32_t L, uint32_t threads);
int curc_fastq_host_parse(const char *fq, uint64_t bytes, uint32_t L, uint32_t threads, char *out,
                          uint64_t out_bytes, uint64_t *n_reads, int *format_error);
/* Record parse + N flag + 2-bit pack of single-end FASTQ bytes on the device
 * (d_fq 16-byte aligned): the parse part of process() and encode
 * (preprocess.cpp:147-199, compress.cu:2151-2180) for one buffer.  Writes
 * d_packed (max_reads * ceil(L/32) words), d_isN (optional; max_reads bytes
 * rounded up to a multiple of 4) and d_seq_off (max_reads u64: where each
 * read's sequence line starts in d_fq); *n_reads receives the record count.
 * CURC_ELIMIT when the bytes hold more than max_reads records. */
int curc_fastq_pack(curc_ctx *ctx, const char *d_fq, uint64_t bytes, uint32_t L, uint64_t max_reads,
                    uint64_t *d_packed, uint8_t *d_isN, uint64_t *d_seq_off, uint64_t *n_reads);
/* No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  curc_fastq_pack with the same arguments and results, but
 * allocating nothing: its counters and tile counts live in a fixed 8 MiB
 * device array that each GPU holds until the process ends.  Inputs of more
 * than CURC_FASTQ_STATIC_MAX_BYTES bytes fail with CURC_ELIMIT.  Calls that
 * overlap on one GPU share that array and race: serialise them. */
#define CURC_FASTQ_STATIC_MAX_BYTES 16000000000ull
int curc_fastq_pack_static(curc_ctx *ctx, const char *d_fq, uint64_t bytes, uint32_t L, uint64_t max_reads,
                           uint64_t *d_packed, uint8_t *d_isN, uint64_t *d_seq_off, uint64_t *n_reads);

/* ---- decompression: read reconstruction -----------------------------------
 * The read rebuild of curc -d (decompress_block, decompress.cu:909-994, and
 * preserve_order_process, :472-577) on the device, from a block's decoded
 * streams (after the entropy decoders and the '%' expansion of the texts,
 * recoverRef, :418-444, which stay on the host).  Every read is a slice of
 * the HQ, LQ or N-read LQ text; a read of the HQ text gets its mismatch
 * symbols (decoder rows for contexts A, T, C, G) at its decoded offsets and,
 * on strand '1', is reverse complemented.  Output lines are L bases and a
 * newline: order not preserved, HQ/mapped records in stream order, then LQ,
 * then N-LQ (pairs regrouped by pe_slot when paired); order preserved, by
 * read id.  Paired output goes to out1 / out2 (n_reads / 2 lines each),
 * single-end output to out1 (n_reads lines).  All pointers are host memory.
 * Inconsistent streams (counts that do not add up, offsets or positions out
 * of range) fail with CURC_EINVAL. */
typedef struct curc_decode_job {
    uint32_t L;
    uint32_t flags;             /* CURC_RUN_PRESERVE_ORDER, CURC_RUN_PAIRED */
    uint64_t n_reads;           /* reads_count of the block */
    const char *hq;             /* texts after the '%' expansion */
    uint64_t hq_len;
    const char *lq;
    uint64_t lq_len;
    const char *nlq;
    uint64_t nlq_len;
    uint64_t n_rec;             /* HQ and mapped records */
    const char *strand;         /* n_rec '0'/'1' */
    const uint8_t *mis_count;   /* n_rec */
    const uint8_t *mis_base;    /* n_mis recoded mismatch symbols (0..3) */
    uint64_t n_mis;
    const uint8_t *mis_table;   /* 16 bytes */
    uint32_t max_mm;            /* floor(L / 5) */
    const uint16_t *const *mis_off; /* stream for m mismatches at [m - 1], m = 1..max_mm */
    const uint64_t *mis_off_len;    /* u16 values in each */
    /* order not preserved */
    const uint16_t *read_off;   /* n_rec position deltas */
    const uint16_t *lq_off;     /* n_lq, L - delta */
    uint64_t n_lq;
    const uint16_t *nlq_off;    /* n_nlq, L - delta */
    uint64_t n_nlq;
    const uint32_t *pe_slot;    /* paired: stream index of the read of each output slot (2k file 1, 2k+1 file 2) */
    /* order preserved */
    const uint64_t *id_pos;     /* n_reads joined-text positions, by read id */
} curc_decode_job;
int curc_decode_block(curc_ctx *ctx, const curc_decode_job *job, char *out1, char *out2);

/* ---- learned candidate scorer ----------------------------------------------
 * No counterpart in the reference.  A 32 -> 128 -> 128 -> 1 network (ReLU,
 * bfloat16 matrix products on the MFMA units, float32 accumulation and
 * master weights) trained with the library's own forward / backward / Adam
 * kernels on mapping outcomes that exact verification produced.  Parameters
 * are float32 in the order W1[32][128], b1[128], W2[128][128], b2[128],
 * w3[128], b3 (CURC_MLP_NPARAMS values).  Sample features are float32
 * [n][32]; labels float32 in {0, 1}; x / y / logit are device pointers. */
#define CURC_MLP_NPARAMS 20865u
typedef struct curc_mlp curc_mlp;
curc_mlp *curc_mlp_create(curc_ctx *ctx);
/* may be called after the context was destroyed */
void curc_mlp_destroy(curc_mlp *m);
int curc_mlp_set_params(curc_mlp *m, const float *params, uint32_t n);  /* also resets Adam */
int curc_mlp_get_params(curc_mlp *m, float *params, uint32_t n);
int curc_mlp_forward(curc_mlp *m, const float *d_x, uint64_t n, float *d_logit);
/* gradients of the mean binary cross-entropy of sigmoid(logit), no update */
int curc_mlp_gradients(curc_mlp *m, const float *d_x, const float *d_y, uint64_t n, float *grads);
/* one Adam step (lr; beta 0.9 / 0.999, eps 1e-8) on the batch; *loss, when
 * given, receives the batch's mean loss before the step */
int curc_mlp_train_step(curc_mlp *m, const float *d_x, const float *d_y, uint64_t n, float lr, float *loss);
/* Learned mapping: with CURC_RUN_LEARNED, the mapping stage checks each
 * unsettled read after 2, 4, 8, 16 and 32 probes of a strand pass; a
 * read whose logit is below its threshold tries no further probes in that
 * pass: `threshold_mapped` for a read that already has a mapping within
 * max_mismatch, `threshold_unmapped` for one that has none yet.  Records
 * stay exact (verification decides); reads may map with more mismatches
 * than the exhaustive loop finds, or not at all.  The network must outlive
 * its use; NULL disables. */
int curc_ctx_set_scorer(curc_ctx *ctx, curc_mlp *m, float threshold_mapped, float threshold_unmapped);
/* Traced mapping (CURC_RUN_TRACE): reads whose slot is a multiple of
 * `every` contribute one record per checkpoint, up to max_records. */
int curc_ctx_set_trace(curc_ctx *ctx, uint32_t every, uint64_t max_records);

/* ---- launch shapes --------------------------------------------------------
 * No counterpart in the reference.  Knobs that change speed only, never
 * results: workgroup caps of the overlap probe, the exact-matching probe and
 * the mapping kernels, keys per thread of a radix-sort tile, the streams
 * curc_batch_run uses when its caller passes 0, the size of the overlap
 * stage's prefix k-mer filter, the list bound below which an
 * exact-matching loop runs fused in one launch, the form of the overlap
 * probe, the mapping index layout and the exact-matching list handling.
 * curc_amd/autotune.py
 * times candidate values on a sample block and sets the fastest. */
typedef struct curc_tuning {
    uint32_t ov_grid;     /* >= 1 (default 16384) */
    uint32_t fm_grid;     /* >= 1 (default 4096) */
    uint32_t map_grid;    /* >= 1 (default 32768) */
    uint32_t radix_items; /* 8, 16 or 32 (default 16) */
    uint32_t workers;     /* 1 .. CURC_BATCH_MAX_WORKERS (default 4) */
    uint32_t filter_bits; /* overlap k-mer filter bits per prefix: 8 .. 128, a power of two (default 32) */
    uint32_t fm_fused;    /* exact-matching lists up to this bound run their offset loop in one
                             single-workgroup launch; 0: never (default 4096) */
    uint32_t ov_queue;    /* overlap probe: 1 queues the suffixes that pass the k-mer filter in
                             LDS and searches their buckets in full batches; 2 as 1, with the
                             filter minimizer-blocked and its verdicts for all offsets computed in
                             one pass beforehand (reads of at most 256 bases; others use 1);
                             3: as 2, each thread taking a quarter of a 64-suffix verdict word
                             against a bitmap of the suffixes linked so far, so that only the
                             suffixes that search their bucket are touched;
                             0: one suffix per thread throughout (default 3) */
    uint32_t map_both;    /* 1: both strand passes of the mapping in one launch, on an index that
                             holds a bucket's forward and reverse-complement slots in one 128-byte
                             line; 2: as 1, on an index of one 32-byte slot per bucket (positions
                             of larger buckets in an overflow array); 0: one launch and one index
                             per strand (default 2) */
    uint32_t fm_static;   /* 1: exact matching keeps its lists in place (linked reads skipped by
                             flag and skip pointer, suffixes screened by a prefix filter); 0: both
                             lists compacted after every offset (default 1) */
    uint32_t fq_host;     /* host threads that parse FASTQ records of host bytes before the copy
                             (curc_fastq_host_parse), so that only the sequences are copied to the
                             device; 0: the bytes are copied as they are and parsed on the device.
                             At most CURC_FQ_HOST_MAX (default 8) */
} curc_tuning;
#define CURC_FQ_HOST_MAX 64u
int curc_ctx_get_tuning(curc_ctx *ctx, curc_tuning *t);
int curc_ctx_set_tuning(curc_ctx *ctx, const curc_tuning *t);

#ifdef __cplusplus
}
#endif
#endif
