// grid.hpp — internal grid-build entry points shared by the launchers and the
// headless pipeline.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbf_params.h"

namespace pbfx {

// Segment compaction of sorted keys into unique keys / offsets / count, with an
// optional per-slot segment index (segOf[s] = index of the cell holding slot s).
// blockCounts/blockOffsets hold compact_scratch_elems(N) entries each.
uint32_t compact_scratch_elems(uint32_t N);
void compact_segments(uint32_t* uniq, uint32_t* offsets, uint32_t* count, uint32_t* segOf,
                      const uint32_t* keysSorted, uint32_t N, uint32_t* blockCounts, uint32_t* blockOffsets,
                      hipStream_t s);

// Slot pairs of the cells after compact_segments (segOf, offsets): each cell's slots two
// at a time, in slot order, pairs[p] = (first, second or ~0 when the cell's count is
// odd); *count = number of pairs.  Same scratch as compact_segments.
void build_slot_pairs(uint2* pairs, uint32_t* count, const uint32_t* segOf, const uint32_t* offsets, uint32_t N,
                      uint32_t* blockCounts, uint32_t* blockOffsets, hipStream_t s);

// Compaction scratch of one stream, for LaunchCellRangesCompact calls on it (reserved
// and released with the stream's drop-in staging).  The reference keeps one scratch per
// process (cell_compact.cu:50-56), which two streams compacting at once would share.
bool reserve_stream_compact_scratch(uint32_t N, hipStream_t s);
void release_stream_compact_scratch(hipStream_t s);

// Fill n words with a value (a kernel, so it is safe inside captured graphs).
void fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s);

}  // namespace pbfx
