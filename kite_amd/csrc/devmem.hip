// devmem.hip -- device memory that grown buffers gave up, freed in one go
// (the reason is in common.hpp, beside retire_device).
#include <mutex>
#include <vector>

#include "common.hpp"

namespace kite {

namespace {

struct Retired {
    void *p;
    size_t bytes;
    int device;
};

std::mutex g_mu;
std::vector<Retired> g_retired;
size_t g_bytes = 0;

// Frees every retired allocation after its device has finished its work.
// Called with g_mu held.
void free_locked() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    int synced = -1;
    for (const Retired &r : g_retired) {
        if (r.device != synced) {
            (void)hipSetDevice(r.device);
            (void)hipDeviceSynchronize();
            synced = r.device;
        }
        (void)hipFree(r.p);
    }
    g_retired.clear();
    g_bytes = 0;
    (void)hipSetDevice(cur);
}

}  // namespace

void retire_device(void *p, size_t bytes) {
    int dev = 0;
    KITE_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(g_mu);
    g_retired.push_back(Retired{p, bytes, dev});
    g_bytes += bytes;
    if (g_bytes > RETIRE_CAP) free_locked();
}

void free_retired() {
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_retired.empty()) free_locked();
}

}  // namespace kite
