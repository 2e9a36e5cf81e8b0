/*
 * Blackframe XPU/HIP — the extension (plugin) ABI entry point of libblackframe_hip.so.
 *
 * The structures, constants and negotiation rules are the reference's, so that its extension
 * loader can open this library like any other XPU backend plugin:
 *   Blackframe/Extensions/Abi/Include/Blackframe/Extensions/ExtensionAbi.h (query, api, status codes)
 *   Blackframe/XPU/BackendAbi/Include/Blackframe/XPU/XpuBackendAbi.h:20-55 (enumeration interface v1)
 * v1 of the XPU interface is enumeration only (device count and descriptors); rendering goes
 * through the blackframe_hip_* entry points of blackframe_hip.h.
 */
#ifndef BLACKFRAME_XPU_EXTENSION_H
#define BLACKFRAME_XPU_EXTENSION_H

#include <stdint.h>

#define BLACKFRAME_EXTENSION_ABI_MAJOR UINT32_C(1)
#define BLACKFRAME_EXTENSION_ABI_MINOR UINT32_C(0)
#define BLACKFRAME_EXTENSION_QUERY_SYMBOL "blackframe_query_extension"

typedef uint32_t blackframe_status;
#define BLACKFRAME_STATUS_SUCCESS ((blackframe_status)0u)
#define BLACKFRAME_STATUS_INVALID_ARGUMENT ((blackframe_status)1u)
#define BLACKFRAME_STATUS_UNSUPPORTED_ABI ((blackframe_status)2u)
#define BLACKFRAME_STATUS_INTERFACE_NOT_FOUND ((blackframe_status)3u)
#define BLACKFRAME_STATUS_INVALID_STATE ((blackframe_status)4u)
#define BLACKFRAME_STATUS_OUT_OF_RANGE ((blackframe_status)5u)
#define BLACKFRAME_STATUS_EXTENSION_FAILURE ((blackframe_status)6u)

typedef struct blackframe_interface_id {
    uint64_t high;
    uint64_t low;
} blackframe_interface_id;

typedef struct blackframe_utf8_view {
    const char* data;
    uint32_t size;
    uint32_t reserved;
} blackframe_utf8_view;

typedef struct blackframe_extension_query_v1 {
    uint32_t struct_size;
    uint32_t abi_major;
    uint32_t abi_minor;
    uint32_t reserved;
} blackframe_extension_query_v1;

typedef blackframe_status (*blackframe_extension_get_interface_fn)(void* extension_context,
                                                                   blackframe_interface_id interface_id,
                                                                   uint32_t requested_major, uint32_t requested_minor,
                                                                   void* out_interface, uint32_t out_interface_size);
typedef void (*blackframe_extension_shutdown_fn)(void* extension_context);

typedef struct blackframe_extension_api_v1 {
    uint32_t struct_size;
    uint32_t abi_major;
    uint32_t abi_minor;
    uint32_t reserved;
    void* extension_context;
    blackframe_extension_get_interface_fn get_interface;
    blackframe_extension_shutdown_fn shutdown;
} blackframe_extension_api_v1;

#define BLACKFRAME_XPU_BACKEND_ABI_MAJOR UINT32_C(1)
#define BLACKFRAME_XPU_BACKEND_ABI_MINOR UINT32_C(0)
#define BLACKFRAME_XPU_BACKEND_INTERFACE_ID_HIGH UINT64_C(0x96c6e70dd06b4b93)
#define BLACKFRAME_XPU_BACKEND_INTERFACE_ID_LOW UINT64_C(0xa9514ab777f49d37)

typedef uint32_t blackframe_xpu_device_kind;
#define BLACKFRAME_XPU_DEVICE_KIND_UNKNOWN ((blackframe_xpu_device_kind)0u)
#define BLACKFRAME_XPU_DEVICE_KIND_CPU ((blackframe_xpu_device_kind)1u)
#define BLACKFRAME_XPU_DEVICE_KIND_GPU ((blackframe_xpu_device_kind)2u)
#define BLACKFRAME_XPU_DEVICE_KIND_ACCELERATOR ((blackframe_xpu_device_kind)3u)

typedef struct blackframe_xpu_device_descriptor_v1 {
    uint32_t struct_size;
    blackframe_xpu_device_kind kind;
    uint64_t identifier_high;
    uint64_t identifier_low;
    blackframe_utf8_view name;
    blackframe_utf8_view vendor;
    uint64_t reserved[4];
} blackframe_xpu_device_descriptor_v1;

typedef blackframe_status (*blackframe_xpu_get_device_count_fn)(void* backend_context, uint32_t* out_device_count);
typedef blackframe_status (*blackframe_xpu_get_device_descriptor_fn)(void* backend_context, uint32_t device_index,
                                                                     blackframe_xpu_device_descriptor_v1* out_descriptor);

typedef struct blackframe_xpu_backend_api_v1 {
    uint32_t struct_size;
    uint32_t abi_major;
    uint32_t abi_minor;
    uint32_t reserved;
    void* backend_context;
    blackframe_utf8_view backend_name;
    blackframe_xpu_get_device_count_fn get_device_count;
    blackframe_xpu_get_device_descriptor_fn get_device_descriptor;
    uint64_t reserved_slots[4];
} blackframe_xpu_backend_api_v1;

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Negotiates the extension ABI: query->struct_size and out_extension->struct_size must cover the
 * v1 structures, query->abi_major must be 1. get_interface serves the XPU backend interface v1
 * (one descriptor per visible HIP device, kind GPU, identifier = the device UUID).
 * Replaces: the per-plugin blackframe_query_extension (ExtensionAbi.h, e.g.
 *           Plugins/XPU/ReferenceCpu/Source/ReferenceCpuExtension.cpp).
 */
blackframe_status blackframe_query_extension(const blackframe_extension_query_v1* query,
                                             blackframe_extension_api_v1* out_extension);

#ifdef __cplusplus
}
#endif

#endif /* BLACKFRAME_XPU_EXTENSION_H */
