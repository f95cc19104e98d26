#include "workspace.h"

#include <map>
#include <mutex>
#include <utility>
#include <vector>

namespace bbhx {

namespace {
struct Buf {
    void* ptr = nullptr;
    size_t bytes = 0;
};
struct StreamSet {
    Buf slot[WS_SLOTS];
};
struct DeviceWs {
    std::map<hipStream_t, StreamSet> sets;
    std::vector<void*> retired;  // replaced buffers, freed by workspace_release_all
};
std::map<int, DeviceWs> g_ws;
std::mutex g_mutex;
}  // namespace

void* workspace_get(int slot, size_t bytes, hipStream_t stream)
{
    if (slot < 0 || slot >= WS_SLOTS) return nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return nullptr;
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceWs& d = g_ws[dev];
    Buf& b = d.sets[stream].slot[slot];
    if (b.bytes >= bytes && b.ptr) return b.ptr;
    // never allocate while the stream is being captured into a graph (an
    // allocation there invalidates the capture): the caller must run the same
    // entry point once on the capture stream, eagerly, before capturing
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (cs != hipStreamCaptureStatusNone) return nullptr;
    // grow by at least 1.5x to avoid repeated reallocation on slowly growing sizes
    size_t want = bytes < 4096 ? 4096 : bytes;
    if (b.bytes && want < b.bytes + b.bytes / 2) want = b.bytes + b.bytes / 2;
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (b.ptr) d.retired.push_back(b.ptr);
    b.ptr = p;
    b.bytes = want;
    return p;
}

void workspace_release_all()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return;
    std::lock_guard<std::mutex> lock(g_mutex);
    auto it = g_ws.find(dev);
    if (it == g_ws.end()) return;
    (void)hipDeviceSynchronize();
    for (auto& kv : it->second.sets)
        for (Buf& b : kv.second.slot)
            if (b.ptr) (void)hipFree(b.ptr);
    for (void* p : it->second.retired) (void)hipFree(p);
    g_ws.erase(it);
}

size_t workspace_bytes()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 0;
    std::lock_guard<std::mutex> lock(g_mutex);
    auto it = g_ws.find(dev);
    if (it == g_ws.end()) return 0;
    size_t n = 0;
    for (auto& kv : it->second.sets)
        for (Buf& b : kv.second.slot) n += b.bytes;
    return n;
}

}  // namespace bbhx
