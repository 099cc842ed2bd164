// radix.h -- stable LSD radix sort of 64-bit keys on the device (radix.hip): per-pass
// count / scan / scatter launches, no device-wide atomics.
#pragma once
#include <cstddef>
#include <cstdint>
#include <hip/hip_runtime.h>

namespace tsne {

constexpr int kRadixBits = 8;  // digit width of one pass

// Scratch bytes a sort of n items needs.
size_t radix_scratch_bytes(int n);

// Sorts bits [begin_bit, end_bit) of keys[0..n), carrying vals; keys_alt / vals_alt are
// buffers of n elements.  On return keys / vals point at the sorted data (the pointers
// are exchanged with the alternates when the last pass wrote there).  0 on success.
int radix_sort_pairs_u64(uint64_t *&keys, uint64_t *&keys_alt, int *&vals, int *&vals_alt, int n,
                         int begin_bit, int end_bit, void *scratch, size_t scratch_bytes,
                         hipStream_t st);

// The same for keys only.
int radix_sort_keys_u64(uint64_t *&keys, uint64_t *&keys_alt, int n, int begin_bit, int end_bit,
                        void *scratch, size_t scratch_bytes, hipStream_t st);

}  // namespace tsne
