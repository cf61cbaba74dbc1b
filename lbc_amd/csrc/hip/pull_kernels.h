// pull_kernels.h -- gather-form fast path for periodic, all-fluid lattices.
//
// The push step (k_bulk) reads each node's own populations and scatters the
// post-collision values to the neighbours; every store is shifted by one node,
// which on MI355X costs ~25% of achievable bandwidth.  For a lattice without
// walls, boundary table or stale edges the same arithmetic can run in gather
// form instead: the buffers then hold POST-collision populations f*, and one
// launch per step gathers f(x) = f*(x - c_k) (streaming.c:21-42 with both axes
// wrapping), forms the lagged moments of the gathered node, collides
// (collision.c:24-37) and stores aligned, non-temporal f* of the next step.
// Rounding is identical to the push path: the same moments_ref / collide_ref
// (FP64) or moments_f32 / collide_f32 (FP32-shifted) run on the same values.
//
// Conversions (bit-exact):
//   to gather form   k_collide_inplace: f*_k = collide(f_k) in the current buffer
//   back to the reference's buffers   k_gather: f_k = gather(f*_{k-1}) into the
//                    buffer that held f*_k; the other buffer keeps f*_{k-1},
//                    which is exactly the reference's f_new after the swap.
#pragma once
#include "step_kernels.h"

namespace lbmx {

constexpr int kPullThreads = 256;

struct PullArgs {
    Geo g;
    StepFlags s;
    Fields F;
    int tiles_per_row;
};

template <typename T>
__device__ __forceinline__ void gather_node(const T *__restrict__ S, const Geo &g, int x, int y,
                                            T f[Q])
{
    const int xm = (x == 0) ? g.Nx - 1 : x - 1, xp = (x == g.Nx - 1) ? 0 : x + 1;
    const int ym = (y == 0) ? g.Ny - 1 : y - 1, yp = (y == g.Ny - 1) ? 0 : y + 1;
    const size_t r0 = (size_t)y * g.pitch, rm = (size_t)ym * g.pitch, rp = (size_t)yp * g.pitch;
#pragma unroll
    for (int k = 0; k < Q; k++) {
        const int cx = D2Q9::cx[k], cy = D2Q9::cy[k];
        const int xs = cx == 1 ? xm : (cx == -1 ? xp : x);        // source x - c_x
        const size_t row = cy == 1 ? rm : (cy == -1 ? rp : r0);   // source y - c_y
        f[k] = S[k * g.plane + row + xs];
    }
}

__device__ __forceinline__ void store_nt(double *p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void store_nt(float *p, float v) { __builtin_nontemporal_store(v, p); }

template <typename T, bool FORCE>
__global__ void __launch_bounds__(kPullThreads) k_pull(const T *__restrict__ S, T *__restrict__ D,
                                                      PullArgs a)
{
    const Geo &g = a.g;
    const int y = blockIdx.x / a.tiles_per_row;
    const int x = (blockIdx.x % a.tiles_per_row) * kPullThreads + threadIdx.x;
    if (x >= g.Nx) return;
    T p[Q];
    gather_node(S, g, x, y, p);
    if constexpr (sizeof(T) == 8) {
        double *f = reinterpret_cast<double *>(p);
        const Moments m = moments_ref(f, a.s.Fx_lag, a.s.Fy_lag);
        collide_ref(f, m, g.om, FORCE, a.s.Fx, a.s.Fy);
    } else if (g.wide) {
        float *q = reinterpret_cast<float *>(p);
        const MomentsW m = moments_wide(q, g.rho0, a.s.Fx_lag, a.s.Fy_lag);
        collide_wide(q, m, g.rho0, g.om, FORCE, a.s.Fx, a.s.Fy);
    } else {
        float *q = reinterpret_cast<float *>(p);
        const float r0 = (float)g.rho0;
        const MomentsF m = moments_f32(q, r0, (float)a.s.Fx_lag, (float)a.s.Fy_lag);
        collide_f32(q, m, r0, (float)g.om, FORCE, (float)a.s.Fx, (float)a.s.Fy);
    }
    const size_t n = (size_t)y * g.pitch + x;
#pragma unroll
    for (int k = 0; k < Q; k++) store_nt(&D[k * g.plane + n], p[k]);
}

// f*_k = collide(f_k) in place (lagged moments: uploaded arrays on the first
// step after an upload, else the node's own moments).
template <typename T>
__global__ void __launch_bounds__(256) k_collide_inplace(T *__restrict__ S, PullArgs a)
{
    const Geo &g = a.g;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= g.Nx) return;
    const size_t n = (size_t)y * g.pitch + x;
    T p[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) p[k] = S[k * g.plane + n];
    if constexpr (sizeof(T) == 8) {
        double *f = reinterpret_cast<double *>(p);
        const Moments m = a.s.lag_arrays ? Moments{a.F.R[n], a.F.U[n], a.F.V[n]}
                                         : moments_ref(f, a.s.Fx_lag, a.s.Fy_lag);
        collide_ref(f, m, g.om, a.s.force, a.s.Fx, a.s.Fy);
    } else if (g.wide) {
        float *q = reinterpret_cast<float *>(p);
        const MomentsW m = a.s.lag_arrays ? MomentsW{a.F.R[n] - g.rho0, a.F.U[n], a.F.V[n]}
                                          : moments_wide(q, g.rho0, a.s.Fx_lag, a.s.Fy_lag);
        collide_wide(q, m, g.rho0, g.om, a.s.force, a.s.Fx, a.s.Fy);
    } else {
        float *q = reinterpret_cast<float *>(p);
        const float r0 = (float)g.rho0;
        MomentsF m;
        if (a.s.lag_arrays) {
            m.dr = (float)(a.F.R[n] - g.rho0);
            m.ux = (float)a.F.U[n];
            m.uy = (float)a.F.V[n];
        } else {
            m = moments_f32(q, r0, (float)a.s.Fx_lag, (float)a.s.Fy_lag);
        }
        collide_f32(q, m, r0, (float)g.om, a.s.force, (float)a.s.Fx, (float)a.s.Fy);
    }
#pragma unroll
    for (int k = 0; k < Q; k++) S[k * g.plane + n] = p[k];
}

// ---- two steps per launch (temporal blocking) ------------------------------------
// A workgroup owns a TX x TY tile of destinations.  Step 1 gathers and collides the
// tile plus a one-node ring (periodic wrap) from S = f*_s in global memory into
// LDS; step 2 gathers the tile from LDS, collides and stores f*_{s+2}.  Each pass
// over the lattice then reads f*_s once (plus the rings, mostly from L2) and
// writes f*_{s+2} once: about 60 % of the HBM traffic of two k_pull launches for
// the same two steps.  The arithmetic is k_pull's, on the same values, so the
// result is bit-identical to two k_pull launches.  The intermediate f*_{s+1} is
// never stored, so a run of steps ends with a k_pull step (the buffer pair then
// holds f*_{k-1} and f*_k, as lbmx_ensure_push needs).
template <typename T, bool W>
__device__ __forceinline__ void pull_collide(T p[Q], const Geo &g, double Fx_lag, double Fy_lag,
                                             bool force, double Fx, double Fy)
{
    if constexpr (sizeof(T) == 8) {
        double *f = reinterpret_cast<double *>(p);
        const Moments m = moments_ref(f, Fx_lag, Fy_lag);
        collide_ref(f, m, g.om, force, Fx, Fy);
    } else if constexpr (W) {
        float *q = reinterpret_cast<float *>(p);
        const MomentsW m = moments_wide(q, g.rho0, Fx_lag, Fy_lag);
        collide_wide(q, m, g.rho0, g.om, force, Fx, Fy);
    } else {
        float *q = reinterpret_cast<float *>(p);
        const float r0 = (float)g.rho0;
        const MomentsF m = moments_f32(q, r0, (float)Fx_lag, (float)Fy_lag);
        collide_f32(q, m, r0, (float)g.om, force, (float)Fx, (float)Fy);
    }
}

struct Pull2Args {
    Geo g;
    int force;
    double Fx, Fy;            // body force of both steps
    double Fx_lag, Fy_lag;    // force the first step's lagged moments were formed with
    int tiles_x;              // tiles per row of tiles
};

constexpr int kPull2Threads = 256;
constexpr int kPull2TX = 64;

template <typename T, bool W, int TY>
__global__ void __launch_bounds__(kPull2Threads) k_pull2(const T *__restrict__ S,
                                                        T *__restrict__ D, Pull2Args a)
{
    constexpr int RX = kPull2TX + 2, RN = RX * (TY + 2);
    __shared__ T buf[Q][RN];
    const Geo &g = a.g;
    const int x0 = (blockIdx.x % a.tiles_x) * kPull2TX;
    const int y0 = (blockIdx.x / a.tiles_x) * TY;

    // step 1: the tile and its ring, from global memory
    for (int r = threadIdx.x; r < RN; r += kPull2Threads) {
        int x = x0 - 1 + r % RX, y = y0 - 1 + r / RX;
        x = (x % g.Nx + g.Nx) % g.Nx;
        y = (y % g.Ny + g.Ny) % g.Ny;
        T p[Q];
        gather_node(S, g, x, y, p);
        pull_collide<T, W>(p, g, a.Fx_lag, a.Fy_lag, a.force, a.Fx, a.Fy);
#pragma unroll
        for (int k = 0; k < Q; k++) buf[k][r] = p[k];
    }
    __syncthreads();

    // step 2: the tile, from LDS
    for (int q = threadIdx.x; q < kPull2TX * TY; q += kPull2Threads) {
        const int tx = q % kPull2TX, ty = q / kPull2TX;
        const int x = x0 + tx, y = y0 + ty;
        if (x >= g.Nx || y >= g.Ny) continue;
        T p[Q];
#pragma unroll
        for (int k = 0; k < Q; k++)
            p[k] = buf[k][(ty + 1 - D2Q9::cy[k]) * RX + (tx + 1 - D2Q9::cx[k])];
        pull_collide<T, W>(p, g, a.Fx, a.Fy, a.force, a.Fx, a.Fy);
        const size_t n = (size_t)y * g.pitch + x;
#pragma unroll
        for (int k = 0; k < Q; k++) store_nt(&D[k * g.plane + n], p[k]);
    }
}

// Three steps per launch: step 1 fills LDS with the tile and a two-node ring, step 2
// collides the tile and a one-node ring from it into registers, which replace the
// LDS contents once every lane has read them, and step 3 is k_pull2's second step.
// LDS holds 9 x 68 x (TY + 4) values.
template <typename T, bool W, int TY>
__global__ void __launch_bounds__(kPull2Threads) k_pull3(const T *__restrict__ S,
                                                        T *__restrict__ D, Pull2Args a)
{
    constexpr int AX = kPull2TX + 4, AN = AX * (TY + 4);
    constexpr int BX = kPull2TX + 2, BN = BX * (TY + 2);
    constexpr int NB = (BN + kPull2Threads - 1) / kPull2Threads;
    __shared__ T buf[Q * AN];
    const Geo &g = a.g;
    const int x0 = (blockIdx.x % a.tiles_x) * kPull2TX;
    const int y0 = (blockIdx.x / a.tiles_x) * TY;

    // step 1: tile + 2-node ring from global memory
    for (int r = threadIdx.x; r < AN; r += kPull2Threads) {
        int x = x0 - 2 + r % AX, y = y0 - 2 + r / AX;
        x = (x % g.Nx + g.Nx) % g.Nx;
        y = (y % g.Ny + g.Ny) % g.Ny;
        T p[Q];
        gather_node(S, g, x, y, p);
        pull_collide<T, W>(p, g, a.Fx_lag, a.Fy_lag, a.force, a.Fx, a.Fy);
#pragma unroll
        for (int k = 0; k < Q; k++) buf[k * AN + r] = p[k];
    }
    __syncthreads();

    // step 2: tile + 1-node ring, from LDS into registers
    T pb[NB][Q];
#pragma unroll
    for (int j = 0; j < NB; j++) {
        const int r = threadIdx.x + j * kPull2Threads;
        if (r < BN) {
            const int bx = r % BX, by = r / BX;   // A coordinates: (bx + 1, by + 1)
#pragma unroll
            for (int k = 0; k < Q; k++)
                pb[j][k] = buf[k * AN + (by + 1 - D2Q9::cy[k]) * AX + (bx + 1 - D2Q9::cx[k])];
            pull_collide<T, W>(pb[j], g, a.Fx, a.Fy, a.force, a.Fx, a.Fy);
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NB; j++) {
        const int r = threadIdx.x + j * kPull2Threads;
        if (r < BN)
#pragma unroll
            for (int k = 0; k < Q; k++) buf[k * BN + r] = pb[j][k];
    }
    __syncthreads();

    // step 3: the tile, from LDS
    for (int q = threadIdx.x; q < kPull2TX * TY; q += kPull2Threads) {
        const int tx = q % kPull2TX, ty = q / kPull2TX;
        const int x = x0 + tx, y = y0 + ty;
        if (x >= g.Nx || y >= g.Ny) continue;
        T p[Q];
#pragma unroll
        for (int k = 0; k < Q; k++)
            p[k] = buf[k * BN + (ty + 1 - D2Q9::cy[k]) * BX + (tx + 1 - D2Q9::cx[k])];
        pull_collide<T, W>(p, g, a.Fx, a.Fy, a.force, a.Fx, a.Fy);
        const size_t n = (size_t)y * g.pitch + x;
#pragma unroll
        for (int k = 0; k < Q; k++) store_nt(&D[k * g.plane + n], p[k]);
    }
}

// f_k = gather(f*_{k-1}) (streaming only).
template <typename T>
__global__ void __launch_bounds__(256) k_gather(const T *__restrict__ S, T *__restrict__ D, Geo g)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= g.Nx) return;
    T p[Q];
    gather_node(S, g, x, y, p);
    const size_t n = (size_t)y * g.pitch + x;
#pragma unroll
    for (int k = 0; k < Q; k++) D[k * g.plane + n] = p[k];
}

}  // namespace lbmx
