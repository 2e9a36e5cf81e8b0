/*
 * The XPU extension entry point: enumeration of the visible HIP devices behind the reference's
 * plugin ABI (include/blackframe_xpu_extension.h).
 */
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/blackframe_xpu_extension.h"

namespace {

const char kBackendName[] = "Blackframe HIP wavefront (MI355X)";
const char kVendor[] = "AMD";

struct device_entry {
    std::string name;
    uint64_t high = 0, low = 0;
};

std::mutex g_mutex;
std::vector<device_entry> g_devices; /* names stay valid until shutdown */
bool g_enumerated = false;

blackframe_utf8_view view(const char* text, size_t size) { return blackframe_utf8_view{text, (uint32_t)size, 0u}; }

void enumerate_locked() {
    if (g_enumerated) return;
    g_enumerated = true;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 0) {
        (void)hipGetLastError();
        count = 0;
    }
    for (int d = 0; d < count; ++d) {
        device_entry e;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) == hipSuccess) {
            e.name = std::string(prop.name) + " (" + prop.gcnArchName + ")";
        } else {
            e.name = "HIP device " + std::to_string(d);
        }
        hipUUID uuid;
        if (hipDeviceGetUuid(&uuid, d) == hipSuccess) {
            std::memcpy(&e.high, uuid.bytes, 8);
            std::memcpy(&e.low, uuid.bytes + 8, 8);
        } else {
            e.low = (uint64_t)d;
        }
        g_devices.push_back(std::move(e));
    }
    (void)hipGetLastError();
}

blackframe_status get_device_count(void*, uint32_t* out) {
    if (!out) return BLACKFRAME_STATUS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(g_mutex);
    enumerate_locked();
    *out = (uint32_t)g_devices.size();
    return BLACKFRAME_STATUS_SUCCESS;
}

blackframe_status get_device_descriptor(void*, uint32_t index, blackframe_xpu_device_descriptor_v1* out) {
    if (!out || out->struct_size < sizeof(blackframe_xpu_device_descriptor_v1)) return BLACKFRAME_STATUS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(g_mutex);
    enumerate_locked();
    if (index >= g_devices.size()) return BLACKFRAME_STATUS_OUT_OF_RANGE;
    const device_entry& e = g_devices[index];
    blackframe_xpu_device_descriptor_v1 d{};
    d.struct_size = sizeof d;
    d.kind = BLACKFRAME_XPU_DEVICE_KIND_GPU;
    d.identifier_high = e.high;
    d.identifier_low = e.low;
    d.name = view(e.name.c_str(), e.name.size());
    d.vendor = view(kVendor, sizeof kVendor - 1);
    *out = d;
    return BLACKFRAME_STATUS_SUCCESS;
}

blackframe_status get_interface(void*, blackframe_interface_id id, uint32_t major, uint32_t, void* out, uint32_t size) {
    if (!out || size < sizeof(blackframe_xpu_backend_api_v1)) return BLACKFRAME_STATUS_INVALID_ARGUMENT;
    if (id.high != BLACKFRAME_XPU_BACKEND_INTERFACE_ID_HIGH || id.low != BLACKFRAME_XPU_BACKEND_INTERFACE_ID_LOW)
        return BLACKFRAME_STATUS_INTERFACE_NOT_FOUND;
    if (major != BLACKFRAME_XPU_BACKEND_ABI_MAJOR) return BLACKFRAME_STATUS_UNSUPPORTED_ABI;
    blackframe_xpu_backend_api_v1 api{};
    api.struct_size = sizeof api;
    api.abi_major = BLACKFRAME_XPU_BACKEND_ABI_MAJOR;
    api.abi_minor = BLACKFRAME_XPU_BACKEND_ABI_MINOR;
    api.backend_name = view(kBackendName, sizeof kBackendName - 1);
    api.get_device_count = &get_device_count;
    api.get_device_descriptor = &get_device_descriptor;
    *static_cast<blackframe_xpu_backend_api_v1*>(out) = api;
    return BLACKFRAME_STATUS_SUCCESS;
}

void shutdown(void*) {
    std::lock_guard<std::mutex> lock(g_mutex);
    g_devices.clear();
    g_enumerated = false;
}

} // namespace

extern "C" blackframe_status blackframe_query_extension(const blackframe_extension_query_v1* query,
                                                        blackframe_extension_api_v1* out) {
    if (!query || !out || query->struct_size < sizeof(blackframe_extension_query_v1) ||
        out->struct_size < sizeof(blackframe_extension_api_v1))
        return BLACKFRAME_STATUS_INVALID_ARGUMENT;
    if (query->abi_major != BLACKFRAME_EXTENSION_ABI_MAJOR) return BLACKFRAME_STATUS_UNSUPPORTED_ABI;
    blackframe_extension_api_v1 api{};
    api.struct_size = sizeof api;
    api.abi_major = BLACKFRAME_EXTENSION_ABI_MAJOR;
    api.abi_minor = BLACKFRAME_EXTENSION_ABI_MINOR;
    api.get_interface = &get_interface;
    api.shutdown = &shutdown;
    *out = api;
    return BLACKFRAME_STATUS_SUCCESS;
}
