// Result copies into user memory (rgl_graph_get_result_data).
//
// One core copies host memory at ~25 GB/s and a device-to-pageable hipMemcpy runs at ~14 GB/s, while the PCIe link
// and host memory sustain several times that. Large copies are therefore split across a few worker threads, and
// device results go through pinned chunks: the transfer of one chunk over the bus overlaps the host copy of the
// previous one into the user's buffer.
#pragma once

#include <cstddef>

namespace rgl {

// Copies `bytes` of host memory; blocks until done.
void hostCopy(void* dst, const void* src, size_t bytes);

// Copies `bytes` of device memory `src` (already complete: the caller has waited for its producer) to host `dst`,
// which may be pageable; blocks until done.
void deviceToHost(void* dst, const void* src, size_t bytes);

} // namespace rgl
