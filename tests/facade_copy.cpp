// facade_copy.cpp -- the device-to-host copy facade_demo uses to read the
// renderer's buffers, in its own translation unit so that the HIP runtime's
// vector types (float3, make_float3) stay out of the facade program's scope.
#include <cstddef>
#include <stdexcept>

#include <hip/hip_runtime_api.h>

void copy_from_device(void* dst, const void* src, std::size_t bytes) {
    if (bytes && hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("hipMemcpy from a renderer buffer failed");
}
