#include "workspace.h"

#include <mutex>

namespace bbhx {

namespace {
constexpr int MAX_DEVICES = 64;
struct Buf {
    void* ptr = nullptr;
    size_t bytes = 0;
};
Buf g_bufs[MAX_DEVICES][WS_SLOTS];
std::mutex g_mutex;
}  // namespace

void* workspace_get(int slot, size_t bytes)
{
    if (slot < 0 || slot >= WS_SLOTS) return nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_mutex);
    Buf& b = g_bufs[dev][slot];
    if (b.bytes >= bytes && b.ptr) return b.ptr;
    // grow by at least 1.5x to avoid repeated reallocation on slowly growing sizes
    size_t want = bytes < 4096 ? 4096 : bytes;
    if (b.bytes && want < b.bytes + b.bytes / 2) want = b.bytes + b.bytes / 2;
    if (b.ptr) {
        // the old buffer may still be read by queued kernels
        if (hipDeviceSynchronize() != hipSuccess) return nullptr;
        (void)hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    b.ptr = p;
    b.bytes = want;
    return p;
}

void workspace_release_all()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return;
    std::lock_guard<std::mutex> lock(g_mutex);
    (void)hipDeviceSynchronize();
    for (int s = 0; s < WS_SLOTS; s++) {
        if (g_bufs[dev][s].ptr) (void)hipFree(g_bufs[dev][s].ptr);
        g_bufs[dev][s] = Buf{};
    }
}

}  // namespace bbhx
