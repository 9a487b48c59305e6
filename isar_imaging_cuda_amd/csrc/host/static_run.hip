// static_run.hip — isar_run_device_static: a full imaging run whose workspace is a file-scope
// device array of the library, so the call allocates no device memory. For callers that cannot
// allocate: code under stream capture, tight serving loops, hosts that forbid hipMalloc.
//
// Each GPU holds its own copy of the array (kStaticWorkspaceBytes) from the library's load until
// the process ends. Host code never passes the array's name to a kernel or a copy: the name is a
// host-side stand-in, and the device address of each GPU's copy comes from hipGetSymbolAddress,
// once per device with that device current. Every call on a GPU uses that GPU's array and leaves
// its contents behind; no kernel of a run reads workspace it has not written in the same run, so
// nothing needs clearing. Calls that overlap on one GPU (two host threads, or two streams whose
// runs are in flight together) race on the array.
#include <cstring>
#include <mutex>
#include <string>

#include "pipeline.hpp"

namespace {

constexpr size_t kStaticWorkspaceBytes = size_t{64} << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kStaticWorkspaceBytes];

std::once_flag g_once[kMaxDevices];
void* g_scratch[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

// the device address of the current device's copy of `scratch`, looked up once per device
void* device_scratch(int* err_dev) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        *err_dev = dev;
        return nullptr;
    }
    std::call_once(g_once[dev], [dev] {
        void* ptr = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        g_scratch[dev] = g_scratch_err[dev] == hipSuccess ? ptr : nullptr;
#ifdef ISAR_EMU
        // CPU emulation: start from all-ones bytes, as emulated device allocations do
        if (g_scratch[dev]) std::memset(g_scratch[dev], 0xFF, kStaticWorkspaceBytes);
#endif
    });
    *err_dev = dev;
    return g_scratch[dev];
}

}  // namespace

extern "C" {

ISAR_API size_t isar_static_workspace_bytes(void) { return kStaticWorkspaceBytes; }

ISAR_API int isar_run_device_static(const isar_radar_params* p, int frames, const void* d_echo,
                                    const double* d_velocity, const void* d_phi0, float* d_img,
                                    const isar_options* opt, void* stream) {
    using namespace isar;
    if (!p || !opt) return fail(ISAR_ERR_INVALID, "null radar parameters or options");
    if (opt->stage_mode != ISAR_RUN_FULL) return fail(ISAR_ERR_INVALID, "the static-workspace run takes full runs only");
    if (frames < 1) return fail(ISAR_ERR_INVALID, "frames must be >= 1");
    int rc = validate_params(*p);
    if (rc) return rc;
    const size_t need = workspace_bytes(*p, frames);
    if (need > kStaticWorkspaceBytes)
        return fail(ISAR_ERR_INVALID, std::to_string(frames) + " frames of " + std::to_string(p->echo_num) + " x " +
                                          std::to_string(p->range_num) + " need " + std::to_string(need) +
                                          " bytes of workspace; the static workspace holds " +
                                          std::to_string(kStaticWorkspaceBytes));
    int dev = -1;
    void* base = device_scratch(&dev);
    if (!base)
        return fail(ISAR_ERR_DEVICE, "static workspace of device " + std::to_string(dev) + ": " +
                                         (dev >= 0 && dev < kMaxDevices ? hipGetErrorString(g_scratch_err[dev])
                                                                        : "device index out of range"));
    const hipStream_t s = static_cast<hipStream_t>(stream);
    Context c;   // host-side state only: every device buffer is carved from `base`
    rc = c.init_in(*p, frames, base, s);
    if (rc) return rc;
    return c.run(frames, d_echo, d_velocity, d_phi0, d_img, *opt, s);
}

}  // extern "C"
