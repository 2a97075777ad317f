// Stable LSD radix sort of (key, u32 value) pairs on the low `end_bit` bits
// of the key, for the point order (Morton / Hilbert keys, src/gpu/mod.rs:
// 422-520 sorts on the host), the stratified insertion order and the
// per-cloud output rows.  8-bit digits; per pass one upsweep (per-tile digit
// counts), one scan (one workgroup per digit) and one downsweep (stable
// ranks from wave64 ballots, scatter).  Equal keys keep their input order.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace wtet {

// temp bytes the sort of n pairs needs (the counts and one spare copy of the
// keys and values)
size_t radix_temp_bytes_u64(uint32_t n, int end_bit);
size_t radix_temp_bytes_u32(uint32_t n, int end_bit);

// keys_in / vals_in are read only and must not overlap the outputs or
// temp; the sorted pairs land in keys_out / vals_out.  Bits at or above
// end_bit are ignored.  Launches on stream s; returns hipErrorInvalidValue
// (and launches nothing) when temp is smaller than the *_temp_bytes size.
hipError_t radix_sort_u64(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out,
                          uint32_t* vals_out, uint32_t n, int end_bit, void* temp,
                          size_t temp_bytes, hipStream_t s);
hipError_t radix_sort_u32(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                          uint32_t* vals_out, uint32_t n, int end_bit, void* temp,
                          size_t temp_bytes, hipStream_t s);

}  // namespace wtet
