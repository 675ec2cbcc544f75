// hostlink.hpp -- device-to-host copies on SDMA engines chosen here.
//
// hipMemcpyAsync leaves the choice of copy engine to the runtime, copy by
// copy.  With a host-to-device stream and a device-to-host stream busy at
// once it can put both directions on one SDMA engine, which then carries
// them one after the other: the two directions share one direction's rate
// (about 57 GB/s on MI355X's PCIe 5.0 x16 instead of about 100 GB/s), and
// the choice can stick for seconds.  HostLink queues device-to-host copies
// through the HSA runtime on engines picked from those the runtime reports
// as preferred for that direction and not for the other one, each copy on
// one of them that has no copy queued, so the runtime's host-to-device
// copies keep an engine of their own.
//
// A copy's source must be device memory whose writes are complete and
// visible to the system (after a HIP event recorded behind them has been
// waited for on the host), its destination page-locked host memory.
#pragma once

#include <hsa/hsa.h>
#include <hsa/hsa_ext_amd.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace kite {

// One device-to-host copy of a HostLink request.
struct CopyPiece {
    void *dst;
    const void *src;
    size_t bytes;
};

class HostLink {
public:
    // The SDMA engines of HIP device `device` (its HSA agent is found by PCI
    // address) and `n_done` completion signals.  usable() is false when no
    // engine qualifies; copies must then take hipMemcpyAsync.
    HostLink(int device, int n_done);
    ~HostLink();
    HostLink(const HostLink &) = delete;
    HostLink &operator=(const HostLink &) = delete;

    bool usable() const { return !engines_.empty(); }
    // Engine bits (hsa_amd_sdma_engine_id_t) for device-to-host copies, in
    // the order tried.
    const std::vector<uint32_t> &engines() const { return engines_; }

    // Queues `bytes` from device memory `src` to page-locked host memory
    // `dst`; completion signal `slot` drops to 0 when the copy has ended.
    // Returns the engine bit used, 0 when the engine refused the copy and
    // the runtime placed it.  Not thread safe: one caller at a time.
    uint32_t d2h(int slot, void *dst, const void *src, size_t bytes);
    // The same for several pieces on one engine and one completion signal,
    // which drops to 0 once all of them (empty ones skipped) have ended.
    uint32_t d2h_many(int slot, const CopyPiece *pieces, int n);
    // Blocks until the copy last queued on `slot` has ended (at once when
    // none is in flight); throws kite::Error when it reported an error.
    void wait(int slot) const;
    // wait() on every slot, reporting no error (teardown after a failure).
    void quiesce() const;

private:
    bool hsa_up_ = false;
    hsa_agent_t gpu_{}, cpu_{};
    std::vector<uint32_t> engines_;
    std::vector<hsa_signal_t> done_;
    uint32_t last_ = 0;
};

}  // namespace kite
