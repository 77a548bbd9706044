/*
 * CPU oracle for the CURC pseudogenome hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation in curc_amd/ is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never links or calls it.
 *
 * It is a single-threaded restatement of the reference's algorithm
 * (src/compress.cu, src/util.cuh, src/preprocess.cpp of BioinfoSZU/CURC).
 * Where the reference's GPU code is racy or leaves ties to an unstable sort,
 * the oracle fixes one deterministic outcome:
 *   - dedup sort ties (compress.cu:1400-1407) are broken by ascending read id;
 *   - overlap claims (compress.cu:1518) are resolved as if suffixes were
 *     processed sequentially in list order;
 *   - mapping-index buckets (compress.cu:369-406) hold the first 12
 *     occurrences by ascending position, stored in descending position;
 *   - N-read sort ties (compress.cu:1179) are broken by ascending N index.
 *
 * The OpenMP build (liboracle_omp.so) parallelises only loops whose
 * iterations are independent, and must produce identical bytes.
 */
#ifndef CURC_ORACLE_H
#define CURC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    uint32_t L;        /* read length */
    uint32_t W;        /* 64-bit words per packed read = ceil(L/32) */
    uint32_t max_off;  /* floor(65 L / 100) */
    uint32_t k;        /* overlap k-mer size */
    uint32_t kip;      /* k-mer index position = L - k - max_off */
    uint32_t max_mm;   /* floor(L / 5) */
} orc_params;

/* Stage selectors for orc_block_run: the run stops after this stage. */
enum {
    ORC_STAGE_DEDUP = 1,
    ORC_STAGE_OVERLAP = 2,
    ORC_STAGE_PRUNE = 3,
    ORC_STAGE_FULLMATCH = 4,
    ORC_STAGE_ASSEMBLY = 5,
    ORC_STAGE_MAP = 6,
    ORC_STAGE_LQ = 7,
    ORC_STAGE_NLQ = 8,
    ORC_STAGE_ALL = 8
};

/* Result fields readable through orc_block_get. */
enum {
    ORC_F_PACKED = 0,        /* u64[n_reads * W] */
    ORC_F_NONN_IDS = 1,      /* u32 */
    ORC_F_N_IDS = 2,         /* u32 */
    ORC_F_SORTED_IDS = 3,    /* u32, non-N ids in dedup order */
    ORC_F_NEXT_DEDUP = 4,    /* u32[n_reads] after duplicate linking */
    ORC_F_NEXT_OVERLAP = 5,  /* u32[n_reads] after the greedy overlap loop */
    ORC_F_PREV_OVERLAP = 6,
    ORC_F_OFFSET_OVERLAP = 7,/* u16[n_reads] */
    ORC_F_NEXT = 8,          /* final (after full match) */
    ORC_F_PREV = 9,
    ORC_F_OFFSET = 10,
    ORC_F_HQ_IDS = 11,
    ORC_F_UNMATCHED_IDS = 12,
    ORC_F_HQ_PG = 13,        /* char */
    ORC_F_HQ_REC = 14,       /* orc_rec */
    ORC_F_MAP_REC = 15,
    ORC_F_MAP_MIS = 16,      /* u8 */
    ORC_F_UNMAPPED = 17,     /* u32 global ids */
    ORC_F_UNMAPPED_N = 18,   /* u32 N indices */
    ORC_F_LQ_PG = 19,
    ORC_F_LQ_REC = 20,
    ORC_F_NLQ_PG = 21,
    ORC_F_NLQ_REC = 22,
    ORC_F_PREV_DEDUP = 23,
    ORC_F_NEXT_PRUNE = 24,
    ORC_F_PREV_PRUNE = 25,
    ORC_F_HQ_PG2 = 26,       /* u64 2-bit PG, ceil(len/32)+1 words */
    ORC_F_LQ_NEXT = 27,      /* u32[n_unmapped], local ids */
    ORC_F_LQ_PREV = 28,
    ORC_F_LQ_OFFSET = 29,
    /* pre-entropy streams of one block (orc_block_streams) */
    ORC_F_S_STRAND = 40,     /* char '0'/'1' per sorted record */
    ORC_F_S_READ_OFF = 41,   /* u16 per sorted record (order not preserved) */
    ORC_F_S_MIS_COUNT = 42,  /* u8 per sorted record */
    ORC_F_S_MIS_BASE = 43,   /* u8 recoded mismatch symbol per mismatch */
    ORC_F_S_MIS_TABLE = 44,  /* 16 bytes: decoder rows for contexts A, T, C, G */
    ORC_F_S_LQ_OFF = 45,     /* u16 per LQ record */
    ORC_F_S_NLQ_OFF = 46,    /* u16 per N-LQ record */
    ORC_F_S_PE_ORDER = 47,   /* u32 global ids in emission order (paired, not preserved) */
    ORC_F_S_ID_POS = 48,     /* u64 per read: joined-PG position (order preserved) */
    ORC_F_S_SORTED_REC = 49, /* orc_rec, HQ+mapped records after the record sort */
    ORC_F_S_MIS_OFF = 64,    /* + m (1..max_mm): u16 mismatch-offset stream for m mismatches */
    ORC_F_COUNT = 64 + 512
};

/* 16-byte alignment record, layout-compatible with the reference's
 * AlignmentRecord (compress.cu:68-77): word = isN | strand << 1 | pos << 2. */
typedef struct orc_rec {
    uint32_t read_id;
    uint32_t pad;
    uint64_t word;
} orc_rec;

typedef struct orc_block orc_block;

int orc_make_params(uint32_t L, orc_params *p);

/* bit helpers (util.cuh) */
uint64_t orc_murmur64(uint64_t x);
uint32_t orc_hash_size(uint64_t item_count);
uint64_t orc_extract_word(const uint64_t *bits, uint64_t start, uint64_t off, uint32_t n);
uint64_t orc_extract_word_unaligned(const uint64_t *bits, uint64_t base_off, uint32_t n);
uint64_t orc_revcomp_kmer(uint64_t kmer, uint32_t k);
int orc_mismatch_words(uint64_t a, uint64_t b);
int orc_forward_mismatch(const uint64_t *ref, uint64_t ref_off, const uint64_t *read,
                         uint32_t L, uint32_t bound);
int orc_rc_mismatch(const uint64_t *ref, uint64_t ref_off, const uint64_t *read,
                    uint32_t L, uint32_t bound);
void orc_pack(const char *ascii, uint64_t n, uint32_t L, uint64_t *packed);

/* Record loop of preprocess.cpp:147-193 over one block's FASTQ bytes (fq2
 * for paired input, ids 2k / 2k+1): reads in id order into out (cap_reads
 * reads of L bytes at most), count in *n_reads.  0 on success; -1 a
 * sequence line of another length than L, -2 a truncated record, -3 paired
 * files with different record counts, -4 more than cap_reads reads. */
int orc_parse_fastq(const char *fq1, uint64_t bytes1, const char *fq2, uint64_t bytes2, uint32_t L, char *out,
                    uint64_t cap_reads, uint64_t *n_reads);

/* Whole-block run over n_reads ASCII reads of length L (block-global ids
 * 0..n_reads-1 in input order).  Returns NULL on invalid input. */
orc_block *orc_block_run(const char *ascii, uint64_t n_reads, uint32_t L, int stop_after);
/* Returns the element count of a field and sets *ptr to its data (owned by
 * the block). */
uint64_t orc_block_get(orc_block *b, int field, const void **ptr);
/* Record sort and stream builders of block_compress (compress.cu:1730-2044)
 * for a block run to ORC_STAGE_NLQ; results via the ORC_F_S_* fields. */
int orc_block_streams(orc_block *b, int preserve_order, int paired_end);
void orc_block_free(orc_block *b);

/* RefMatcher (compress.cu:597-1001, refmatch.cpp): matches the reverse
 * complement of dest against a k=36 index of src (the HQ pseudogenome),
 * rewrites dest with '%' markers and builds the MapOff / MapLen streams.
 * dest_contain_N selects the character path (compress.cu:809-853) over the
 * binary one (:746-807); dest_is_ref marks the self match of src. */
typedef struct orc_refmatch orc_refmatch;
enum {
    ORC_RM_TEXT = 0,      /* char: rewritten destination text */
    ORC_RM_MAP_OFF = 1,   /* bytes: u32 (u64 if src_len > 2^32-1) source positions */
    ORC_RM_MAP_LEN = 2,   /* bytes: byte-frugal 50, then byte-frugal (length - 50) */
    ORC_RM_FOUND = 3,     /* u64 triples (src, len, dest) in collection order */
    ORC_RM_MATCHES = 4,   /* u64 triples after transform, sort and unique */
    ORC_RM_CAPACITY = 5   /* u64 */
};
uint32_t orc_marush36(const char *s);
orc_refmatch *orc_ref_match(const char *src, uint64_t src_len, const char *dest, uint64_t dest_len,
                            int dest_contain_N, int dest_is_ref);
uint64_t orc_refmatch_get(orc_refmatch *r, int field, const void **ptr);
void orc_refmatch_free(orc_refmatch *r);

#ifdef __cplusplus
}
#endif
#endif
