// spk_batch_async.h -- the batched path's steps that do not wait for the
// device, for the serving loop (spk_serve.hip).  The C ABI in spk_hip.h is
// built from them and waits at each call.
#pragma once

#include <cstdint>

#include "spk_hip.h"

namespace spk {

// Admits a replica; its set-up and initial energy are queued on the batch
// stream (after a block in flight) and nothing waits.
int batch_add_async(spk_hip_batch *b, const spk_hip_replica_desc *r, int *id);
// Queues one block with the launch shape already set; returns at once.
int batch_launch(spk_hip_batch *b, int max_sweeps, bool use_graph);
// Waits for everything queued on the batch stream; applies the per-sweep
// rows of the block in flight, if any, and the initial energies of
// replicas admitted meanwhile.
int batch_complete(spk_hip_batch *b);
bool batch_inflight(const spk_hip_batch *b);
// Queues a replica's checksum after everything on the batch stream; the
// value is readable once batch_complete (or any wait) has passed it.
int batch_queue_checksum(spk_hip_batch *b, int id);
uint64_t batch_checksum_value(const spk_hip_batch *b, int id);
// Sweeps replica `id` has left (as of the last completed block).
int batch_left(const spk_hip_batch *b, int id);

}  // namespace spk
