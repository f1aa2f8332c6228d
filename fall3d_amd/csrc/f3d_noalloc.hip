// cmass3d over device-resident arrays with no allocation: the budget of
// cmass3d.f90:24-58 (the cmass3d_ drop-in's kernels) for callers that cannot
// call hipMalloc -- code under HIP graph capture, serving loops, harnesses
// that forbid allocation.  The partial sums live in one file-scope device
// array; the caller passes inputs and outputs.
#include <mutex>

#include "f3d_ctx.h"

namespace {

using namespace f3d;

// members per call, and the scratch they need: 4 quantities x the fixed
// per-member partial count of the reduction (red::kMemberPartials)
constexpr int kMaxMembers = 64;
constexpr size_t kScratchDoubles = 4u * (size_t)red::kMemberPartials * kMaxMembers;  // 512 KiB

// Every call on a device shares this array (each device holds its own
// copy); it keeps the last call's partials.  In host code its name is a
// host-side stand-in, not a device address: only the address that
// hipGetSymbolAddress returns for the current device is ever used.
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};

// the current device's address of `scratch`, looked up once per device
int scratch_ptr(double **out)
{
    int dev = -1;
    F3D_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("f3d_cmass3d_noalloc: device %d beyond the %d supported", dev, kMaxDevices);
        return F3D_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("f3d_cmass3d_noalloc: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
            return F3D_EHIP;
        }
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return F3D_OK;
}

// out[r] = (out[r] * dx) * dy, the scaling cmass3d_ applies on the host
// (cmass3d.f90:191-194), in the same association
__global__ void k_scale_area(double *out, int n, double dx, double dy)
{
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t < n) out[t] = out[t] * dx * dy;
}

}  // namespace

extern "C" int f3d_cmass3d_noalloc(const double *c, const double *cload, const double *cdry,
                                   const double *cwet, int nx, int ny, int nz, int nc,
                                   int nmembers, double dx, double dy, const double *dz,
                                   double *out, void *stream)
{
    if (!c || !cload || !cdry || !cwet || !dz || !out) {
        set_error("f3d_cmass3d_noalloc: null array");
        return F3D_EINVAL;
    }
    if (nx < 2 || ny < 2 || nz < 2 || nc < 1) {
        set_error("f3d_cmass3d_noalloc: nx, ny, nz must be >= 2 and nc >= 1 (got %d %d %d %d)",
                  nx, ny, nz, nc);
        return F3D_EINVAL;
    }
    if (nmembers < 1 || nmembers > kMaxMembers) {
        set_error("f3d_cmass3d_noalloc: %d members; its fixed scratch holds 1 to %d", nmembers,
                  kMaxMembers);
        return F3D_EINVAL;
    }
    double *scr = nullptr;
    TRY(scratch_ptr(&scr));
    Grid g = make_grid(nx, ny, nz, nc, dx, dy);
    g.nm = nmembers;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // member m: c + m*nc*G, deposits (nx,ny,0:nc) at m*(nc+1)*NXY, slot 0 read
    TRY(red::cmass3d(s, g, c, cload, cdry, cwet, (long)(nc + 1) * g.NXY, dz, scr, out));
    const int n = 4 * nmembers;
    hipLaunchKernelGGL(k_scale_area, dim3((n + 255) / 256), dim3(256), 0, s, out, n, dx, dy);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}
