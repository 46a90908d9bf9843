// d3d_capi.hip -- library-level C-ABI entry points: version, device setup and
// the thread-local error string.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "d3d.h"
#include "d3d_runtime.hpp"

namespace d3d {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

}  // namespace d3d

extern "C" int d3d_abi_version(void) { return D3D_ABI_VERSION; }

extern "C" const char *d3d_last_error(void) { return d3d::g_last_error.c_str(); }

extern "C" int d3d_init(int device) {
    d3d::clear_error();
    int count = 0;
    D3D_CHECK_HIP(hipGetDeviceCount(&count));
    D3D_CHECK_ARG(device >= 0 && device < count, "device index out of range");
    D3D_CHECK_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    D3D_CHECK_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        d3d::set_error(std::string("libd3d_hip is built for gfx950, device is ") + prop.gcnArchName);
        return D3D_E_INVALID;
    }
    return d3d::noalloc_scratch(nullptr);
}
