// world_internal.h -- what the server in this library uses of a world beyond
// the C ABI of include/vsim.h.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/vsim.h"

namespace vsim {

// Enqueues copies of scene `id`'s node positions (3 planes of its node
// count) into `pos` on the world's stream, after the work enqueued so far.
// `pos` must be pinned host memory and stay valid until an event recorded on
// the world after this call (vs_world_record) has completed.
int world_read_positions_async(vs_world *w, uint32_t id, float *pos);

// The world's compute stream, and the (timing) event vs_world_record
// returned in `slot`; null for a bad slot.
hipStream_t world_stream(vs_world *w);
hipEvent_t world_event(vs_world *w, uint32_t slot);

}  // namespace vsim
