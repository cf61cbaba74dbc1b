// totals_kernels.h -- FP64 global mass and kinetic energy of a lattice's moments.
//
// mass = sum rho, energy = sum rho (u^2 + v^2) / 2 over the fluid nodes, from the
// FP64 moment arrays compute_macroscopic_fields leaves (src/core/
// compute_macroscopic_fields.c:19-48: NaN on solid nodes, which are skipped by
// the solid rule, so a fluid NaN still shows in the totals).  The reference's
// drivers form such full-grid reductions on the host every step
// (cases/poiseuille/poiseuille.c:126-137); here they report how far a
// reduced-precision run has drifted from an FP64 one.
//
// Fixed-order reduction, so the totals depend only on the field and its size:
// block b of G = min(Ny, kTotMaxBlocks) sums rows b, b + G, ...; each thread a
// fixed column stride of them, then a pairwise tree in LDS; one block then sums
// the G partials the same way.  The partials (2 G doubles) are the only scratch.
#pragma once
#include "batch_kernels.h"

namespace lbmx {

constexpr int kTotThreads = 256;
constexpr int kTotMaxBlocks = 1024;

struct TotalsArgs {
    const double *R, *U, *V;   // moments, rows `pitch` elements apart
    int Nx, Ny, pitch;
    const uint8_t *cls;        // node classes (same pitch, 0 = solid), or null
    const int *mask;           // int solid mask [Ny][Nx] (reference layout), or null
    const BatchLat *lat;       // batch lattice whose wall rule marks the solids, or null
    double *partials;          // 2 x gridDim.x
};

__device__ __forceinline__ void tot_tree(double m, double e, double *out2)
{
    __shared__ double sm[kTotThreads], se[kTotThreads];
    sm[threadIdx.x] = m;
    se[threadIdx.x] = e;
    __syncthreads();
    for (int h = kTotThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sm[threadIdx.x] += sm[threadIdx.x + h];
            se[threadIdx.x] += se[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out2[0] = sm[0];
        out2[1] = se[0];
    }
}

static __global__ void __launch_bounds__(kTotThreads) k_totals_part(TotalsArgs a)
{
    double m = 0.0, e = 0.0;
    for (int y = blockIdx.x; y < a.Ny; y += gridDim.x) {
        const size_t row = (size_t)y * a.pitch;
        for (int x = threadIdx.x; x < a.Nx; x += kTotThreads) {
            bool solid = false;
            if (a.cls) solid = a.cls[row + x] == 0;
            if (a.mask) solid = solid || a.mask[(size_t)y * a.Nx + x] != 0;
            if (a.lat) solid = solid || b_solid(*a.lat, x, y);
            if (solid) continue;
            const double r = a.R[row + x], u = a.U[row + x], v = a.V[row + x];
            m += r;
            e += 0.5 * r * (u * u + v * v);
        }
    }
    tot_tree(m, e, a.partials + 2 * blockIdx.x);
}

static __global__ void __launch_bounds__(kTotThreads) k_totals_final(const double *partials, int n,
                                                                     double *out)
{
    double m = 0.0, e = 0.0;
    for (int b = threadIdx.x; b < n; b += kTotThreads) {
        m += partials[2 * b];
        e += partials[2 * b + 1];
    }
    tot_tree(m, e, out);
}

// Blocks of the first pass for a lattice of Ny rows.
__host__ __device__ __forceinline__ int totals_blocks(int Ny)
{
    return Ny < kTotMaxBlocks ? Ny : kTotMaxBlocks;
}

// Both passes on `stream`; `partials` holds 2 * totals_blocks(Ny) doubles and
// `out` receives {mass, energy}.
static inline hipError_t launch_totals(TotalsArgs a, double *partials, double *out,
                                       hipStream_t stream)
{
    const int G = totals_blocks(a.Ny);
    a.partials = partials;
    hipLaunchKernelGGL(k_totals_part, dim3(G), dim3(kTotThreads), 0, stream, a);
    hipLaunchKernelGGL(k_totals_final, dim3(1), dim3(kTotThreads), 0, stream,
                       (const double *)partials, G, out);
    return hipGetLastError();
}

}  // namespace lbmx
