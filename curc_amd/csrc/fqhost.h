// FASTQ record parse on the host: the record loop of preprocess.cpp:147-199
// (four lines per record, the sequence line kept, its length checked
// against the read length, :152-154), run by several threads over one
// file's bytes.  Only the sequence lines cross PCIe afterwards, half of the
// FASTQ bytes for 150-base reads.
//
// The split is exact: thread t starts at the first line of its byte range
// that looks like a record start ('@', then a sequence line of L bases, then
// a line starting with '+'), thread t-1 follows the record chain (every
// fourth newline) from its own start until it reaches a record start at or
// past that guess, and the guess stands only if the chain lands on it.  If
// any guess fails, the file is parsed again in one pass from its first byte.
#pragma once
#include <cstdint>
#include <vector>

namespace curc {

// Format errors of one file, in the order the device parser reports them.
enum FqHostErr : int {
    FQH_OK = 0,
    FQH_START = 1,  // the bytes do not start with '@'
    FQH_TRUNC = 2,  // the bytes end inside a record (no final newline, or a line count not a multiple of 4)
    FQH_LEN = 3,    // a sequence line does not hold exactly L bytes ("read length isn't fix")
};

// `count` sequences of L bytes each, at byte `off` of the output buffer.
struct FqSeg {
    uint64_t off;
    uint64_t count;
};

struct FqHostParse {
    uint64_t n = 0;            // records
    int err = FQH_OK;
    bool resync = false;       // a thread's start guess failed: parsed again in one pass
    std::vector<FqSeg> segs;   // the sequences in record order
};

// Output bytes fq_host_parse may write for nb bytes of FASTQ, read length L
// and T threads (each record of a valid sequence line takes at least L + 4
// bytes of input).
uint64_t fq_host_cap(uint64_t nb, uint32_t L, uint32_t T);

// Parses the nb bytes at B with T threads (T >= 1) into `out`
// (fq_host_cap(nb, L, T) bytes).
void fq_host_parse(const uint8_t* B, uint64_t nb, uint32_t L, uint32_t T, uint8_t* out, FqHostParse* r);

}  // namespace curc
