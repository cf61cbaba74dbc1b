// noalloc.hip -- lbmx_totals_noalloc: the FP64 mass / kinetic-energy reduction of
// lbmx_totals over caller-owned device arrays, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops).
//
// The reduction's only scratch -- the per-block partial sums of its first pass,
// 2 x kTotMaxBlocks doubles (16 KiB) -- is the file-scope device array below.  Its
// size covers every lattice: the first pass uses min(Ny, kTotMaxBlocks) blocks.
// Host code never passes the array's name to a kernel (in host code the name is a
// stand-in, not a device address); each device's address is looked up once with
// hipGetSymbolAddress, with that device current, and kept per device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "lattice_impl.h"
#include "totals_kernels.h"

using namespace lbmx;

static __device__ __align__(256) double scratch[2 * kTotMaxBlocks];
static_assert(sizeof(scratch) == LBMX_NOALLOC_SCRATCH_BYTES, "scratch size stated in lbmx.h");

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices];   // scratch's address on each device (null: not looked up)
}  // namespace

// scratch's device address on the current device, looked up on its first use there.
static int scratch_of_current_device(double **out)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return lbmx_fail(LBMX_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return lbmx_fail(LBMX_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                             hipGetErrorString(e));
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return LBMX_OK;
}

extern "C" int lbmx_totals_noalloc(int Nx, int Ny, const double *rho, const double *u,
                                   const double *v, const int *mask, double *out, void *stream)
{
    if (!rho || !u || !v || !out) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (Nx < 1 || Ny < 1 || (long long)Nx * Ny > LBMX_TOTALS_NOALLOC_MAX_NODES)
        return lbmx_fail(LBMX_ERR_ARG,
                         "lattice %d x %d: lbmx_totals_noalloc takes 1 .. %lld nodes", Nx, Ny,
                         (long long)LBMX_TOTALS_NOALLOC_MAX_NODES);
    double *partials = nullptr;
    if (int rc = scratch_of_current_device(&partials)) return rc;
    TotalsArgs a{};
    a.R = rho;
    a.U = u;
    a.V = v;
    a.Nx = Nx;
    a.Ny = Ny;
    a.pitch = Nx;
    a.mask = mask;
    // totals_blocks(Ny) <= kTotMaxBlocks blocks write 2 partials each: inside scratch
    HIPCHK(launch_totals(a, partials, out, static_cast<hipStream_t>(stream)));
    return LBMX_OK;
}
