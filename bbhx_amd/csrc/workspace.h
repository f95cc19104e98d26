// Per-device, grow-only scratch buffers owned by the library.
//
// Kernels that need per-(binary, mode) coefficient records or per-binary
// descriptor tables take them from here. A buffer only grows; growing waits
// for the device to go idle before releasing the old one, so in steady state
// (same or smaller problem sizes) no allocation happens on any call.
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
// or nullptr on allocation failure.
void* workspace_get(int slot, size_t bytes);

// Releases every buffer of the current device (after a device synchronize).
void workspace_release_all();

}  // namespace bbhx
