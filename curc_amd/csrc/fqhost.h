// FASTQ record parse on the host: the record loop of preprocess.cpp:147-199
// (four lines per record, the sequence line kept, its length checked
// against the read length, :152-154), run by several threads over one
// file's bytes.  Only the sequence lines cross PCIe afterwards, half of the
// FASTQ bytes for 150-base reads.
//
// The split is exact: range q starts at the first line of its bytes that
// looks like a record start ('@', then a sequence line of L bases, then a
// line starting with '+'), range q-1 follows the record chain (every fourth
// newline) from its own start until it reaches a record start at or past
// that guess, and the guess stands only if the chain lands on it.  If a guess
// fails, the bytes after the last confirmed record start are parsed again in
// one pass.
#pragma once
#include <cstdint>
#include <functional>
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
    bool resync = false;       // a range's start guess failed: the rest parsed again in one pass
    std::vector<FqSeg> segs;   // the sequences in record order
};

// The bytes are cut into FQH_RANGES ranges per thread, which the threads take
// in order, so that the sequences of the first ranges can be handed on (and
// copied to the device) while the later ones are still parsed.
constexpr uint32_t FQH_RANGES = 4;

// Output bytes fq_host_parse may write for nb bytes of FASTQ, read length L
// and T threads (each record of a valid sequence line takes at least L + 4
// bytes of input).
uint64_t fq_host_cap(uint64_t nb, uint32_t L, uint32_t T);

// Called on the parsing thread, in record order, for each segment as soon as
// it and every range before it are parsed and their starts confirmed; first
// is the read index of the segment's first sequence.  May throw: parsing then
// stops, and the exception is rethrown once the threads have ended.
using FqEmit = std::function<void(const FqSeg& seg, uint64_t first)>;

// Parses the nb bytes at B with T threads (T >= 1) into `out`
// (fq_host_cap(nb, L, T) bytes).  On a format error r->segs is empty
// (segments may have been emitted before the error was known).
void fq_host_parse(const uint8_t* B, uint64_t nb, uint32_t L, uint32_t T, uint8_t* out, FqHostParse* r,
                   const FqEmit& emit = {});

}  // namespace curc
