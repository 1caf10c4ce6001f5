// =============================================================================
// energy.hip — GPU energy and power telemetry through ROCm SMI: the MI355X
// counterpart of the reference's NVML monitor (nvml_monitor.cpp:35-193) and
// of the RAPL reader its EnergyMonitor uses (energy_monitor.cpp:27-164).
//
// librocm_smi64 is loaded at run time (dlopen), so the solve path does not
// depend on it: without it, or when the device exposes no energy counter,
// hsps_energy_open returns HSPS_ERR_UNSUPPORTED.  The HIP device is matched
// to its SMI index by PCI domain / bus / device.  Energy comes from the
// device's accumulating counter (counts x resolution in microjoules), power
// from the current socket power.
// =============================================================================
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <rocm_smi/rocm_smi.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "hsps_internal.hpp"

namespace hsps_dev {
namespace {

struct Smi {
    bool ok = false;
    std::string why;
    decltype(&rsmi_init) init = nullptr;
    decltype(&rsmi_num_monitor_devices) num = nullptr;
    decltype(&rsmi_dev_pci_id_get) pci = nullptr;
    decltype(&rsmi_dev_energy_count_get) energy = nullptr;
    decltype(&rsmi_dev_power_get) power = nullptr;
    decltype(&rsmi_dev_power_cap_get) cap = nullptr;
    decltype(&rsmi_dev_temp_metric_get) temp = nullptr;
    decltype(&rsmi_dev_busy_percent_get) busy = nullptr;
    decltype(&rsmi_dev_memory_usage_get) mem = nullptr;
    decltype(&rsmi_dev_name_get) name = nullptr;
};

template <class F> bool sym(void* h, const char* n, F& f) {
    f = reinterpret_cast<F>(dlsym(h, n));
    return f != nullptr;
}

Smi& smi() {
    static Smi s;
    static std::once_flag once;
    std::call_once(once, [] {
        void* h = dlopen("librocm_smi64.so.1", RTLD_NOW | RTLD_LOCAL);
        if (!h) h = dlopen("/opt/rocm/lib/librocm_smi64.so.1", RTLD_NOW | RTLD_LOCAL);
        if (!h) {
            s.why = "librocm_smi64 not found";
            return;
        }
        const bool all = sym(h, "rsmi_init", s.init) && sym(h, "rsmi_num_monitor_devices", s.num) &&
                         sym(h, "rsmi_dev_pci_id_get", s.pci) && sym(h, "rsmi_dev_energy_count_get", s.energy) &&
                         sym(h, "rsmi_dev_power_get", s.power) && sym(h, "rsmi_dev_power_cap_get", s.cap) &&
                         sym(h, "rsmi_dev_temp_metric_get", s.temp) && sym(h, "rsmi_dev_busy_percent_get", s.busy) &&
                         sym(h, "rsmi_dev_memory_usage_get", s.mem) && sym(h, "rsmi_dev_name_get", s.name);
        if (!all) {
            s.why = "librocm_smi64 lacks an entry point";
            return;
        }
        if (s.init(0) != RSMI_STATUS_SUCCESS) {
            s.why = "rsmi_init failed";
            return;
        }
        s.ok = true;
    });
    return s;
}

// domain [63:32], bus [15:8], device [7:3]; function and partition bits ignored
constexpr uint64_t kBdfMask = 0xFFFFFFFF0000FFF8ull;

}  // namespace
}  // namespace hsps_dev

using namespace hsps_dev;

struct hsps_energy_s {
    uint32_t dv = 0;
    int device = 0;
};

extern "C" {

int hsps_energy_open(int device, hsps_energy* out) {
    return guarded([&] {
        if (!out) throw Error(HSPS_ERR_INVALID, "hsps_energy_open: out is NULL");
        *out = nullptr;
        int count = 0;
        HSPS_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count)
            throw Error(HSPS_ERR_INVALID, "hsps_energy_open: device " + std::to_string(device) + " not present");
        Smi& s = smi();
        if (!s.ok) throw Error(HSPS_ERR_UNSUPPORTED, "GPU energy telemetry unavailable: " + s.why);
        char bus[64] = {0};
        HSPS_HIP(hipDeviceGetPCIBusId(bus, sizeof bus, device));
        unsigned dom = 0, b = 0, d = 0, f = 0;
        if (std::sscanf(bus, "%x:%x:%x.%x", &dom, &b, &d, &f) != 4)
            throw Error(HSPS_ERR_HIP, std::string("hsps_energy_open: unexpected PCI id ") + bus);
        const uint64_t want = ((uint64_t)dom << 32) | ((uint64_t)(b & 0xff) << 8) | ((uint64_t)(d & 0x1f) << 3);
        uint32_t n = 0;
        if (s.num(&n) != RSMI_STATUS_SUCCESS || n == 0)
            throw Error(HSPS_ERR_UNSUPPORTED, "GPU energy telemetry unavailable: no SMI devices");
        for (uint32_t i = 0; i < n; ++i) {
            uint64_t id = 0;
            if (s.pci(i, &id) != RSMI_STATUS_SUCCESS || (id & kBdfMask) != want) continue;
            uint64_t e = 0, ts = 0;
            float res = 0.f;
            if (s.energy(i, &e, &res, &ts) != RSMI_STATUS_SUCCESS)
                throw Error(HSPS_ERR_UNSUPPORTED, "GPU energy telemetry unavailable: no energy counter");
            auto* h = new hsps_energy_s;
            h->dv = i;
            h->device = device;
            *out = h;
            return HSPS_OK;
        }
        throw Error(HSPS_ERR_UNSUPPORTED, std::string("GPU energy telemetry: no SMI device at ") + bus);
    });
}

int hsps_energy_read(hsps_energy E, double* joules, double* watts, double* timestamp_s) {
    return guarded([&] {
        if (!E) throw Error(HSPS_ERR_INVALID, "hsps_energy_read: null handle");
        Smi& s = smi();
        uint64_t e = 0, ts = 0;
        float res = 0.f;
        if (s.energy(E->dv, &e, &res, &ts) != RSMI_STATUS_SUCCESS)
            throw Error(HSPS_ERR_UNSUPPORTED, "rsmi_dev_energy_count_get failed");
        if (joules) *joules = (double)e * (double)res * 1e-6;
        if (timestamp_s) *timestamp_s = (double)ts * 1e-9;
        if (watts) {
            uint64_t p = 0;
            RSMI_POWER_TYPE t = RSMI_INVALID_POWER;
            *watts = s.power(E->dv, &p, &t) == RSMI_STATUS_SUCCESS ? (double)p * 1e-6 : -1.0;
        }
        return HSPS_OK;
    });
}

int hsps_energy_info(hsps_energy E, double* power_cap_w, double* temperature_c, double* busy_pct,
                     double* vram_used_mb, char* name, int name_len) {
    return guarded([&] {
        if (!E) throw Error(HSPS_ERR_INVALID, "hsps_energy_info: null handle");
        Smi& s = smi();
        if (power_cap_w) {
            uint64_t c = 0;
            *power_cap_w = s.cap(E->dv, 0, &c) == RSMI_STATUS_SUCCESS ? (double)c * 1e-6 : -1.0;
        }
        if (temperature_c) {
            int64_t t = 0;
            *temperature_c = s.temp(E->dv, RSMI_TEMP_TYPE_JUNCTION, RSMI_TEMP_CURRENT, &t) == RSMI_STATUS_SUCCESS
                                 ? (double)t * 1e-3
                                 : -1.0;
        }
        if (busy_pct) {
            uint32_t b = 0;
            *busy_pct = s.busy(E->dv, &b) == RSMI_STATUS_SUCCESS ? (double)b : -1.0;
        }
        if (vram_used_mb) {
            uint64_t u = 0;
            *vram_used_mb = s.mem(E->dv, RSMI_MEM_TYPE_VRAM, &u) == RSMI_STATUS_SUCCESS ? (double)u / 1048576.0 : -1.0;
        }
        if (name && name_len > 0) {
            name[0] = 0;
            if (s.name(E->dv, name, (size_t)name_len) != RSMI_STATUS_SUCCESS) name[0] = 0;
        }
        return HSPS_OK;
    });
}

int hsps_energy_close(hsps_energy E) {
    delete E;
    return HSPS_OK;
}

}  // extern "C"
