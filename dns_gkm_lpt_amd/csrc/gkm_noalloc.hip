// gkm_noalloc.hip -- gkm_step_noalloc: one Eulerian step on caller-owned device
// buffers with no device allocation (declared in include/dnsgkm.h).
//
// The step's scratch -- the flux-difference accumulator dW of the x and y
// sweeps (5 nc^3 doubles) and the device dt the sweeps read -- lives in the
// file-scope array `scratch` below, sized for nc <= kNoallocMaxNc.  Each GPU
// the library runs on holds its own copy until the process ends; its address
// on a device is taken once, with that device current, by hipGetSymbolAddress
// and kept per device: host code uses that address, never the array's name.
// Every call on a GPU shares that GPU's array and overwrites all of it that
// the step reads before reading it (the x sweep writes every dW entry, the dt
// kernel the dt), so nothing needs clearing; calls that overlap on one GPU
// (different streams, or threads) race on it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/dnsgkm.h"
#include "gkm_config.hpp"
#include "gkm_kernels.hpp"

namespace {

constexpr int kNoallocMaxNc = 64;
constexpr long kNoallocDW = 5L * kNoallocMaxNc * kNoallocMaxNc * kNoallocMaxNc;  // dW doubles
constexpr long kNoallocDoubles = kNoallocDW + 32;                                // + dt (aligned)

__device__ __align__(256) double scratch[kNoallocDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device
int g_slots[kMaxDevices] = {};        // concurrent sweep workgroups per device

__global__ void set_dt_kernel(double* p, double dt)
{
    if (threadIdx.x == 0) *p = dt;
}

int put_msg(char* msg, int len, int code, const std::string& m)
{
    if (msg && len > 0) std::snprintf(msg, (size_t)len, "%s", m.c_str());
    return code;
}

// the device's scratch address and sweep slots, set up on the first call on it
int device_setup(int dev, double** ptr, int* slots, char* msg, int len)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return put_msg(msg, len, GKM_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        int cus = 0;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return put_msg(msg, len, GKM_EHIP, std::string("hipDeviceGetAttribute: ") + hipGetErrorString(e));
        g_slots[dev] = 2 * (cus > 0 ? cus : 1);  // the sweep kernels run two workgroups per CU
        g_scratch[dev] = static_cast<double*>(p);
    }
    *ptr = g_scratch[dev];
    *slots = g_slots[dev];
    return GKM_OK;
}

}  // namespace

extern "C" int gkm_step_noalloc(const gkm_config* cfg, const double* W, double* Wout, double dt, void* stream,
                                char* msg, int msglen)
{
    if (!cfg || !W || !Wout) return put_msg(msg, msglen, GKM_EINVAL, "null argument");
    if (W == Wout) return put_msg(msg, msglen, GKM_EINVAL, "W and Wout must be different buffers");
    {
        const std::string m = gkm::config_error(*cfg, gkm::sweep_grid_ok);
        if (!m.empty()) return put_msg(msg, msglen, GKM_EINVAL, m);
    }
    if (cfg->nc > kNoallocMaxNc)
        return put_msg(msg, msglen, GKM_EINVAL,
                       "nc = " + std::to_string(cfg->nc) + " exceeds the no-allocation maximum of " +
                           std::to_string(kNoallocMaxNc) + " (scratch of 5 x 64^3 doubles per GPU)");
    if (!(dt > 0.0) || !std::isfinite(dt)) return put_msg(msg, msglen, GKM_EINVAL, "dt must be positive and finite");
    if (cfg->device < 0 || cfg->device >= kMaxDevices) return put_msg(msg, msglen, GKM_EINVAL, "device out of range");
    hipError_t e = hipSetDevice(cfg->device);
    if (e != hipSuccess) return put_msg(msg, msglen, GKM_EHIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    double* sc = nullptr;
    int slots = 0;
    int rc = device_setup(cfg->device, &sc, &slots, msg, msglen);
    if (rc) return rc;

    const hipStream_t s = static_cast<hipStream_t>(stream);
    const bool mixed = cfg->precision == GKM_MIXED;
    double* ddt = sc + kNoallocDW;
    hipLaunchKernelGGL(set_dt_kernel, dim3(1), dim3(64), 0, s, ddt, dt);
    e = hipGetLastError();
    if (e != hipSuccess) return put_msg(msg, msglen, GKM_EHIP, std::string("dt kernel: ") + hipGetErrorString(e));
    const int nc = cfg->nc;
    for (int dir = 0; dir < 3; dir++) {
        gkm::SweepArgs a{};
        a.W = W;
        a.dW = sc;
        a.Wout = Wout;
        a.cmp_in = nullptr;  // the plain FP64 update in both precisions
        a.cmp_out = nullptr;
        a.dt = ddt;
        a.plane = (long)nc * nc * nc;
        a.nc = nc;
        a.dx = gkm::config_dx(*cfg);
        a.inv_dx = 1.0 / a.dx;
        a.g = gkm::config_gas(*cfg);
        a.tile = gkm::sweep_first_tile(dir, nc);  // a new context's tiles
        a.seg = gkm::sweep_seg_len(dir, nc, a.tile, slots);
        a.nseg = gkm::sweep_nseg(nc, a.seg);
        a.mode = dir;  // x: dW = dF, y: dW += dF, z: Wout = W - (dW + dF)/dx
        e = gkm::launch_sweep(dir, a, mixed, s);
        if (e != hipSuccess) return put_msg(msg, msglen, GKM_EHIP, std::string("sweep launch: ") + hipGetErrorString(e));
    }
    return put_msg(msg, msglen, GKM_OK, "");
}
