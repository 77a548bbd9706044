"""A literal restatement of the reference's FASTQ loop (preprocess.cpp:72-243)
for the tests: byte-by-byte chunk cut, record parse per chunk, block close
rule, ids.  It follows the reference statement by statement, with no search
shortcuts, so that curc_amd.fastq (which jumps between candidates) is checked
against the loop as written.  Only used on small inputs."""


def process_blocks(data1, data2=None, block_ratio=1.0, io_chunk=50_000_000, reset_between_candidates=False):
    """Returns the blocks as lists of read sequences (bytes) in block id
    order, plus the (start, end) byte range of each block.
    reset_between_candidates=True is a planted mistake (the run counter
    reset at every candidate) that the tests expect to be caught."""
    paired = data2 is not None
    fastq_bytes = len(data1)
    if paired and len(data2) != fastq_bytes:
        raise ValueError("paired-end file size isn't equal")
    # compress(): read_len from the first record's sequence line
    first_nl = data1.index(b"\n")
    read_len = data1.index(b"\n", first_nl + 1) - first_nl - 1
    if not paired:
        block_bytes = int(float(fastq_bytes) * block_ratio)
    else:
        block_bytes = int(2 * float(fastq_bytes) * block_ratio)
    io_buffer_size = io_chunk
    blocks, ranges = [], []
    reads = []
    total_read_bytes = 0
    last_total_read_bytes = 0
    while total_read_bytes < fastq_bytes:
        if not paired:
            size = min(io_buffer_size, fastq_bytes - total_read_bytes)
        else:
            size = min(io_buffer_size // 2, fastq_bytes - total_read_bytes)
        io_buffer = data1[total_read_bytes:total_read_bytes + size]
        io_buffer_2 = data2[total_read_bytes:total_read_bytes + size] if paired else None
        buffer_end = len(io_buffer)
        total_read_bytes += buffer_end
        if total_read_bytes < fastq_bytes:
            first_line_len = 0
            if io_buffer[0] != ord("@"):
                raise ValueError("io_buffer doesn't start with @")
            while io_buffer[first_line_len] != ord("\n"):
                first_line_len += 1
            new_buffer_end = buffer_end - first_line_len
            equal_count = 0
            found = False
            while new_buffer_end > 0:
                if io_buffer[new_buffer_end - 1] == ord("\n") and io_buffer[new_buffer_end] == ord("@"):
                    if reset_between_candidates:
                        equal_count = 0
                    k = new_buffer_end
                    for j in range(first_line_len):
                        if io_buffer[j] == io_buffer[k]:
                            equal_count += 1
                            if equal_count > 5:
                                found = True
                                break
                        else:
                            equal_count = 0
                        k += 1
                    if found:
                        break
                new_buffer_end -= 1
            if not found:
                raise ValueError("no cut (the reference would not advance)")
            total_read_bytes -= buffer_end - new_buffer_end
            buffer_end = new_buffer_end
        i = j = 0
        while i < buffer_end:
            while io_buffer[i] != ord("\n"):
                i += 1
            i += 1
            seq_start = i
            while io_buffer[i] != ord("\n"):
                i += 1
            if read_len != i - seq_start:
                raise ValueError("read length isn't fix")
            reads.append(bytes(io_buffer[seq_start:i]))
            i += 1
            while io_buffer[i] != ord("\n"):
                i += 1
            i += 1
            while io_buffer[i] != ord("\n"):
                i += 1
            i += 1
            if paired:
                while io_buffer_2[j] != ord("\n"):
                    j += 1
                j += 1
                seq_start = j
                while io_buffer_2[j] != ord("\n"):
                    j += 1
                if read_len != j - seq_start:
                    raise ValueError("read length isn't fix")
                reads.append(bytes(io_buffer_2[seq_start:j]))
                j += 1
                while io_buffer_2[j] != ord("\n"):
                    j += 1
                j += 1
                while io_buffer_2[j] != ord("\n"):
                    j += 1
                j += 1
        block_read_bytes = total_read_bytes - last_total_read_bytes
        if paired:
            block_read_bytes *= 2
        if block_read_bytes > block_bytes:
            blocks.append(reads)
            ranges.append((last_total_read_bytes, total_read_bytes))
            reads = []
            last_total_read_bytes = total_read_bytes
    if total_read_bytes - last_total_read_bytes > 0:
        blocks.append(reads)
        ranges.append((last_total_read_bytes, total_read_bytes))
    return blocks, ranges
