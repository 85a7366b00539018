// radix_sort.hpp — stable LSD radix sort of (uint32 key, uint32 value) pairs.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace pbfx {

// Bytes of temporary storage needed to sort N pairs over key bits [beginBit, endBit).
size_t radix_sort_temp_bytes(uint32_t N, int beginBit, int endBit);

// Sorts (kin, vin) into (kout, vout); vin == nullptr means values 0..N-1.
// kin/vin are read only; kout/vout must not alias them.  Stream ordered,
// no allocation, no host synchronisation (safe under graph capture).
// Returns false (and records an error) when the temp buffer is too small.
// items: keys per thread of a tile (4, 8 or 16; 0 picks by N).  digitBits: the widest
// digit of a pass (1..8; 0 means 8): the bits are spread evenly over
// ceil(bits / digitBits) passes.
bool radix_sort_pairs(void* temp, size_t tempBytes, const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
                      uint32_t* vout, uint32_t N, int beginBit, int endBit, hipStream_t s, int items = 0,
                      int digitBits = 0);

// Number of key bits needed to hold keys in [0, numKeys).
inline int key_bits_for(uint64_t numKeys) {
    int b = 0;
    while (b < 32 && (1ull << b) < numKeys) ++b;
    return b;
}

}  // namespace pbfx
