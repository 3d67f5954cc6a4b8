// Library-internal access to a tp_context for the serving loop (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/toothpaste_amd.h"

namespace tp {

// Sets the message returned by tp_last_error on this thread.
int set_error(int code, const char* msg);
hipStream_t context_stream(tp_context* c);
void context_params(tp_context* c, tp_params* p);
// Non-zero when the blind-rotation and LWE key-switching keys are imported.
int context_ready(tp_context* c);
int bind_device(tp_context* c);
// Advances the slots active[0..count) by up to `chunk_steps` blind-rotation steps; slots that reach
// the last step write their extracted LWE to ext[slot].  All pointers are device pointers.
int launch_chunk(tp_context* c, size_t count, const uint32_t* active, uint32_t* step, uint64_t* state,
                 const uint64_t* lwes, const uint64_t* tv, const uint32_t* tv_index, uint64_t* ext,
                 uint32_t chunk_steps);
int context_device(tp_context* c);
// Full bootstrap of count ciphertexts: blind rotation + sample extraction into ext ([count][N + 1]
// device words, fully overwritten), then the LWE key switch of ext into out.
int bootstrap_into(tp_context* c, size_t count, const uint64_t* lwes, const uint64_t* tv, size_t tv_stride,
                   uint64_t* ext, uint64_t* out);
// Resolves, for `device` (the current device), the device address of the scratch array behind
// tp_bootstrap_batch_device_noalloc (noalloc.hip); idempotent, thread-safe.
int noalloc_setup(int device);
// Key-switches rows[0..count) of `in` (stride in_stride) into out[0..count).
int launch_ks_rows(tp_context* c, size_t count, const uint64_t* in, size_t in_stride, const uint32_t* rows,
                   uint64_t* out);

}  // namespace tp
