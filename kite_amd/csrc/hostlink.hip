// hostlink.hip -- device-to-host copies on SDMA engines chosen here (the
// engine choice and its reason are in hostlink.hpp).
#include "hostlink.hpp"

#include "common.hpp"

namespace kite {

namespace {

struct AgentSearch {
    uint32_t domain, bus, dev;
    hsa_agent_t gpu{};
    hsa_agent_t cpu{};
    bool found_gpu = false, found_cpu = false;
};

hsa_status_t visit_agent(hsa_agent_t a, void *arg) {
    AgentSearch &s = *static_cast<AgentSearch *>(arg);
    hsa_device_type_t type;
    if (hsa_agent_get_info(a, HSA_AGENT_INFO_DEVICE, &type) != HSA_STATUS_SUCCESS) return HSA_STATUS_SUCCESS;
    if (type == HSA_DEVICE_TYPE_CPU && !s.found_cpu) {
        s.cpu = a;
        s.found_cpu = true;
    }
    if (type == HSA_DEVICE_TYPE_GPU && !s.found_gpu) {
        // LocationID: bus in bits 15:8, device in 7:3, function in 2:0
        uint32_t bdf = 0, domain = 0;
        if (hsa_agent_get_info(a, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_BDFID, &bdf) == HSA_STATUS_SUCCESS &&
            hsa_agent_get_info(a, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_DOMAIN, &domain) == HSA_STATUS_SUCCESS &&
            domain == s.domain && (bdf >> 8) == s.bus && ((bdf >> 3) & 31u) == s.dev) {
            s.gpu = a;
            s.found_gpu = true;
        }
    }
    return HSA_STATUS_SUCCESS;
}

}  // namespace

HostLink::HostLink(int device, int n_done) {
    if (hsa_init() != HSA_STATUS_SUCCESS) return;
    hsa_up_ = true;
    int bus = 0, dev = 0, domain = 0;
    KITE_HIP(hipDeviceGetAttribute(&bus, hipDeviceAttributePciBusId, device));
    KITE_HIP(hipDeviceGetAttribute(&dev, hipDeviceAttributePciDeviceId, device));
    KITE_HIP(hipDeviceGetAttribute(&domain, hipDeviceAttributePciDomainID, device));
    AgentSearch s{(uint32_t)domain, (uint32_t)bus, (uint32_t)dev};
    if (hsa_iterate_agents(visit_agent, &s) != HSA_STATUS_SUCCESS || !s.found_gpu) return;
    gpu_ = s.gpu;
    // host memory is named by the CPU agent nearest the GPU (any CPU agent
    // when the runtime does not say)
    hsa_agent_t near{};
    if (hsa_agent_get_info(gpu_, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_NEAREST_CPU, &near) == HSA_STATUS_SUCCESS &&
        near.handle)
        cpu_ = near;
    else if (s.found_cpu)
        cpu_ = s.cpu;
    else
        return;
    // the engines that serve host traffic are the first NUM_SDMA_ENG ones;
    // those after them are meant for device-to-device xGMI copies
    uint32_t n_host = 0;
    if (hsa_agent_get_info(gpu_, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_NUM_SDMA_ENG, &n_host) != HSA_STATUS_SUCCESS ||
        n_host == 0)
        return;
    const uint32_t host_mask = n_host >= 16 ? 0xFFFFu : ((1u << n_host) - 1u);
    uint32_t pref_d2h = 0, pref_h2d = 0;
    if (hsa_amd_memory_get_preferred_copy_engine(cpu_, gpu_, &pref_d2h) != HSA_STATUS_SUCCESS) pref_d2h = 0;
    if (hsa_amd_memory_get_preferred_copy_engine(gpu_, cpu_, &pref_h2d) != HSA_STATUS_SUCCESS) pref_h2d = 0;
    // host-to-device copies go to the runtime's preferred engines for that
    // direction (engine 0 when it names none): keep off them
    const uint32_t h2d = pref_h2d ? pref_h2d : 1u;
    uint32_t m = pref_d2h & host_mask & ~h2d;
    if (!m) m = host_mask & ~h2d;
    if (!m) return;
    for (int bit = 15; bit >= 0; bit--)
        if (m & (1u << bit)) engines_.push_back(1u << bit);
    done_.resize((size_t)n_done);
    for (hsa_signal_t &d : done_) {
        if (hsa_signal_create(0, 0, nullptr, &d) != HSA_STATUS_SUCCESS) {
            d.handle = 0;
            engines_.clear();
        }
    }
}

HostLink::~HostLink() {
    for (hsa_signal_t d : done_)
        if (d.handle) {
            while (hsa_signal_wait_scacquire(d, HSA_SIGNAL_CONDITION_LT, 1, UINT64_MAX, HSA_WAIT_STATE_BLOCKED) >= 1) {
            }
            (void)hsa_signal_destroy(d);
        }
    if (hsa_up_) (void)hsa_shut_down();
}

uint32_t HostLink::d2h(int slot, void *dst, const void *src, size_t bytes) {
    const CopyPiece p{dst, src, bytes};
    return d2h_many(slot, &p, 1);
}

uint32_t HostLink::d2h_many(int slot, const CopyPiece *pieces, int n) {
    if (!usable()) throw Error("HostLink::d2h: no SDMA engine for device-to-host copies");
    hsa_signal_t d = done_.at((size_t)slot);
    int live = 0;
    for (int i = 0; i < n; i++) live += pieces[i].bytes ? 1 : 0;
    hsa_signal_store_relaxed(d, live);
    if (!live) return 0;
    // an engine of the list with nothing queued, else the one this link
    // used last (its queue holds only this link's own copies)
    uint32_t free_mask = 0;
    if (hsa_amd_memory_copy_engine_status(cpu_, gpu_, &free_mask) != HSA_STATUS_SUCCESS) free_mask = 0;
    uint32_t e = 0;
    for (uint32_t x : engines_)
        if (free_mask & x) {
            e = x;
            break;
        }
    if (!e) e = last_ ? last_ : engines_[0];
    uint32_t used = e;
    for (int i = 0; i < n; i++) {
        const CopyPiece &p = pieces[i];
        if (!p.bytes) continue;
        hsa_status_t st = hsa_amd_memory_async_copy_on_engine(p.dst, cpu_, p.src, gpu_, p.bytes, 0, nullptr, d,
                                                              (hsa_amd_sdma_engine_id_t)e, false);
        if (st != HSA_STATUS_SUCCESS) {
            // the engine refused this copy (the runtime can withhold an
            // engine): the runtime picks one, on the same completion signal
            used = 0;
            st = hsa_amd_memory_async_copy(p.dst, cpu_, p.src, gpu_, p.bytes, 0, nullptr, d);
        }
        if (st != HSA_STATUS_SUCCESS) {
            // the pieces not queued will not count the signal down
            hsa_signal_subtract_relaxed(d, live);
            const char *msg = nullptr;
            (void)hsa_status_string(st, &msg);
            throw Error(std::string("device-to-host copy failed to queue: ") +
                        (msg ? msg : std::to_string((int)st)));
        }
        live--;
    }
    last_ = e;
    return used;
}

void HostLink::wait(int slot) const {
    hsa_signal_t d = done_.at((size_t)slot);
    hsa_signal_value_t v;
    while ((v = hsa_signal_wait_scacquire(d, HSA_SIGNAL_CONDITION_LT, 1, UINT64_MAX, HSA_WAIT_STATE_BLOCKED)) >= 1) {
    }
    if (v < 0) throw Error("device-to-host copy on an SDMA engine reported an error");
}

void HostLink::quiesce() const {
    for (size_t i = 0; i < done_.size(); i++)
        if (done_[i].handle) {
            while (hsa_signal_wait_scacquire(done_[i], HSA_SIGNAL_CONDITION_LT, 1, UINT64_MAX,
                                             HSA_WAIT_STATE_BLOCKED) >= 1) {
            }
        }
}

}  // namespace kite
