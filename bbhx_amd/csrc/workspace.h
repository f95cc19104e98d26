// Scratch buffers owned by the library, one set per (device, stream).
//
// Kernels that need per-(binary, mode) coefficient records or per-binary
// descriptor tables take them from here. Work queued on different streams
// gets different buffers, so two callers driving the library from two streams
// never share scratch memory. A buffer only grows; a buffer that is replaced by
// a larger one is retired, not freed, until workspace_release_all(), so kernels
// still queued on the old buffer, and launch sequences captured in a graph
// before the growth, keep valid memory (no device synchronisation and no free
// on the growth path). A buffer never grows while its stream is being captured
// into a graph: an allocation there would invalidate the capture, so the
// request fails instead, and the caller must first run the same work once,
// eagerly, on the capture stream. In steady state (same or smaller problem
// sizes) no allocation happens on any call.
// Calls that overlap on one stream cannot overlap in time (a stream runs its
// work in order); replays of one graph that run concurrently with each other
// would share its buffers and must not be issued.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>

namespace bbhx {

// Slot ids: each caller uses its own slot so that buffers used by the same
// launch sequence never alias.
enum WsSlot : int {
    WS_PHM_REC = 0,
    WS_RESP_REC = 1,
    WS_TABLE = 2,
    WS_PARTIAL = 3,
    WS_FUSED = 4,
    WS_GRAD = 5,
    WS_GRAD_AUX = 6,
    WS_PREP = 7,
    WS_INTERP = 8,
    WS_SLOTS = 9,
};

// Returns a device pointer with at least 'bytes' bytes on the current device,
// private to work queued on 'stream', or nullptr on allocation failure (also
// when the buffer would have to grow while 'stream' is capturing).
void* workspace_get(int slot, size_t bytes, hipStream_t stream);

// Frees every buffer of the current device, retired ones included (after a
// device synchronize). Graphs captured before the call must not be replayed
// after it.
void workspace_release_all();

// Bytes held on the current device (live and retired buffers).
size_t workspace_bytes();

}  // namespace bbhx
