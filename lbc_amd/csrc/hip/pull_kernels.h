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
// State: the buffer pop[pstar] holds f*_{k-1}, the post-collision field of the
// last step done (the reference's f_new after the driver's swap; a periodic
// all-fluid lattice has no stale entry), and f_k = gather(f*_{k-1}).
// Conversions (bit-exact):
//   to gather form   k_collide_inplace: f*_k = collide(f_k) in the current buffer,
//                    which is the next step (its streaming is the next gather)
//   back to the reference's buffers   k_gather: f_k = gather(f*_{k-1}) into the
//                    other buffer; pop[pstar] stays as the reference's f_new.
#pragma once
#include <utility>
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
// never stored; f*_{s+2} is all a following step or lbmx_ensure_push needs.
// NF: no body force in this launch or in the lagged moments (Fx = Fy = Fx_lag =
// Fy_lag = 0), so the force terms are left out at compile time (bit-identical, see
// moments_nf).
template <typename T, bool W, bool NF = false>
__device__ __forceinline__ void pull_collide(T p[Q], const Geo &g, double Fx_lag, double Fy_lag,
                                             bool force, double Fx, double Fy)
{
    if constexpr (NF) {
        Fx_lag = Fy_lag = Fx = Fy = 0.0;
        force = false;
    }
    if constexpr (sizeof(T) == 8 && NF) {
        double *f = reinterpret_cast<double *>(p);
        collide_nf(f, moments_nf(f), g.om);
    } else if constexpr (sizeof(T) == 8) {
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

template <typename T, bool W, int TY, bool NF = false>
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
        pull_collide<T, W, NF>(p, g, a.Fx_lag, a.Fy_lag, a.force, a.Fx, a.Fy);
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
        pull_collide<T, W, NF>(p, g, a.Fx, a.Fy, a.force, a.Fx, a.Fy);
        const size_t n = (size_t)y * g.pitch + x;
#pragma unroll
        for (int k = 0; k < Q; k++) store_nt(&D[k * g.plane + n], p[k]);
    }
}

// Three steps per launch: step 1 fills LDS with the tile and a two-node ring, step 2
// collides the tile and a one-node ring from it into registers, which replace the
// LDS contents once every lane has read them, and step 3 is k_pull2's second step.
// LDS holds 9 x 68 x (TY + 4) values.
template <typename T, bool W, int TY, bool NF = false>
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
        pull_collide<T, W, NF>(p, g, a.Fx_lag, a.Fy_lag, a.force, a.Fx, a.Fy);
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
            pull_collide<T, W, NF>(pb[j], g, a.Fx, a.Fy, a.force, a.Fx, a.Fy);
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
        pull_collide<T, W, NF>(p, g, a.Fx, a.Fy, a.force, a.Fx, a.Fy);
        const size_t n = (size_t)y * g.pitch + x;
#pragma unroll
        for (int k = 0; k < Q; k++) store_nt(&D[k * g.plane + n], p[k]);
    }
}

// ---- D steps per launch by row streaming -------------------------------------
// A workgroup owns a strip of up to LW - 2(D-1) columns (a.tx: rounded down so
// that a strip's rows are whole 32-byte sectors) and a segment of rows, and
// walks down the segment one row per iteration.  Time level l (1..D) is computed
// by its own LW/64 wavefronts on a row of LW - 2(l-1) columns (the strip plus
// the ring level l still needs); level 1 gathers from S in global memory, levels
// 2..D gather from level l-1's rows in LDS, and level D stores f*_{s+D} to D.
// Level l works two rows behind level l-1, so everything an iteration reads was
// written in an earlier one.  Each level keeps its last R rows in an LDS ring:
// R = 4 needs one barrier per iteration, R = 3 two (the new row replaces the
// oldest once the first barrier says it has been read).  Level 1 keeps P rows of
// loads in flight in registers.  Compared with the tiles of k_pull2/k_pull3, no
// row is computed twice (only the 2(D-1) ring columns of a strip are), and the
// loads of one wavefront overlap the arithmetic of the others.  Same arithmetic
// on the same values as D k_pull launches: bit-identical.
struct StreamArgs {
    Geo g;
    int force;
    double Fx, Fy;            // body force of every step
    double Fx_lag, Fy_lag;    // force the first step's lagged moments were formed with
    int strips;               // strips per row of the lattice
    int tx;                   // output columns per strip (at most LW - 2(D - 1))
    int seg_rows;             // output rows per workgroup
    int segs;                 // row segments
};

// gather_node for a wavefront-uniform row y: the row and plane offsets stay in
// scalar registers and each lane adds one of three fixed 32-bit byte offsets (its
// column x - 1, x, x + 1, wrapped), so the loads need no per-lane address math.
template <typename T>
__device__ __forceinline__ void gather_row(const T *__restrict__ S, const Geo &g, int y,
                                           unsigned bxm, unsigned bx, unsigned bxp, T f[Q])
{
    const int ym = (y == 0) ? g.Ny - 1 : y - 1, yp = (y == g.Ny - 1) ? 0 : y + 1;
    const char *r0 = reinterpret_cast<const char *>(S + (size_t)y * g.pitch);
    const char *rm = reinterpret_cast<const char *>(S + (size_t)ym * g.pitch);
    const char *rp = reinterpret_cast<const char *>(S + (size_t)yp * g.pitch);
#pragma unroll
    for (int k = 0; k < Q; k++) {
        const int cx = D2Q9::cx[k], cy = D2Q9::cy[k];
        const char *base = (cy == 1 ? rm : (cy == -1 ? rp : r0)) + (size_t)k * g.plane * sizeof(T);
        const unsigned off = cx == 1 ? bxm : (cx == -1 ? bxp : bx);
        f[k] = *reinterpret_cast<const T *>(base + off);
    }
}

// LDS rings of k_stream.  Level l + 1 gathers population k of its row r from row
// r - c_y(k) of level l, so of each row a level writes, populations with c_y = -1
// are read while it is the newest row, c_y = 0 one iteration later and c_y = +1
// two iterations later.  With three-row rings (the new row replaces the oldest
// between two barriers) population k therefore keeps only 2 + c_y(k) rows: 18
// values per column instead of 27 (FP64 four levels: 26.8 instead of 40.2 KB of
// LDS, six workgroups per CU instead of four; five levels: 35.1 KB, four per CU).
// Four-row rings (one barrier) keep all four rows of every population: their
// slim form, 3 + c_y(k) rows, made the FP32 eight-level kernel slower (266 vs
// 282 GLUPS at 16384^2, 245 vs 259 at 4096^2).
template <int D, int LW, int R>
struct StreamShape {
    static constexpr int TX = LW - 2 * (D - 1);                  // output columns
    static constexpr int width(int l) { return LW - 2 * (l - 1); }
    static constexpr int rows(int k) { return R == 3 ? 2 + D2Q9::cy[k] : R; }
    static constexpr int first(int k)                            // ring rows before k's
    {
        int o = 0;
        for (int j = 0; j < k; j++) o += rows(j);
        return o;
    }
    static constexpr int per_col = first(Q);
    static constexpr int offset(int l)                           // LDS values before level l
    {
        int o = 0;
        for (int m = 1; m < l; m++) o += per_col * width(m);
        return o;
    }
    static constexpr int lds = offset(D);                        // levels 1..D-1
    // ring row of population k holding lattice row r (r - rbase >= -2(D-1); 12 D is
    // a multiple of every ring length)
    static __device__ __forceinline__ int slot(int k, int r, int rbase)
    {
        return first(k) + (int)((unsigned)(r - rbase + 12 * D) % (unsigned)rows(k));
    }
};

// Level L (2..D) of k_stream.  With a four-row ring (R = 4) a level writes the slot
// of a row nobody reads in the same iteration: one barrier per iteration.  With a
// three-row ring the new row replaces the oldest, which the next level reads in the
// same iteration, so the write waits for a first barrier and a second one publishes
// it.
template <typename T, bool W, int D, int LW, int R, bool NF, int L>
__device__ __forceinline__ void stream_level(T *buf, const T *__restrict__ S, T *__restrict__ Dst,
                                             const StreamArgs &a, int lane, int x0, int y0, int y1,
                                             int iters)
{
    using Sh = StreamShape<D, LW, R>;
    constexpr int Win = Sh::width(L - 1), Wout = Sh::width(L);
    const Geo &g = a.g;
    const int rbase = y0 - (D - 1);
    const T *in = buf + Sh::offset(L - 1);
    for (int t = 0; t < iters; t++) {
        const int r = rbase + t - 2 * (L - 1);
        const bool active = r >= y0 - (D - L) && r <= y1 - 1 + (D - L) && lane < Wout;
        T p[Q];
        if (active) {
#pragma unroll
            for (int k = 0; k < Q; k++)
                p[k] = in[Sh::slot(k, r - D2Q9::cy[k], rbase) * Win + lane + 1 - D2Q9::cx[k]];
            pull_collide<T, W, NF>(p, g, a.Fx, a.Fy, a.force, a.Fx, a.Fy);
            if constexpr (L == D) {
                const int x = x0 + lane;
                if (x < g.Nx && lane < a.tx) {
                    char *row = reinterpret_cast<char *>(Dst + (size_t)r * g.pitch);
                    const unsigned bx = (unsigned)x * (unsigned)sizeof(T);
#pragma unroll
                    for (int k = 0; k < Q; k++)
                        store_nt(reinterpret_cast<T *>(row + (size_t)k * g.plane * sizeof(T) + bx),
                                 p[k]);
                }
            }
        }
        if constexpr (R == 3) __syncthreads();
        if constexpr (L < D) {
            if (active) {
                T *out = buf + Sh::offset(L) + lane;
#pragma unroll
                for (int k = 0; k < Q; k++) out[Sh::slot(k, r, rbase) * Wout] = p[k];
            }
        }
        __syncthreads();
    }
}

template <typename T, bool W, int D, int LW, int R, bool NF, int... Ls>
__device__ __forceinline__ void stream_levels(int level, T *buf, const T *__restrict__ S,
                                              T *__restrict__ Dst, const StreamArgs &a, int lane,
                                              int x0, int y0, int y1, int iters,
                                              std::integer_sequence<int, Ls...>)
{
    ((level == Ls + 2
          ? stream_level<T, W, D, LW, R, NF, Ls + 2>(buf, S, Dst, a, lane, x0, y0, y1, iters)
          : void()),
     ...);
}

template <typename T, bool W, int D, int LW, int P, int R, bool NF = false>
__global__ void __launch_bounds__(LW * D) k_stream(const T *__restrict__ S, T *__restrict__ Dst,
                                                  StreamArgs a)
{
    using Sh = StreamShape<D, LW, R>;
    static_assert(LW % 64 == 0 && Sh::TX > 0 && (R == 3 || R == 4), "strip shape");
    __shared__ T buf[Sh::lds];
    const Geo &g = a.g;
    const int level = __builtin_amdgcn_readfirstlane(threadIdx.x / LW) + 1;
    const int lane = threadIdx.x % LW;
    // XCD-aware order: consecutive workgroups go to the 8 XCDs in turn, so XCD i
    // takes the i-th eighth of the strip-major work list (strips of one segment,
    // then the next segment) and neighbouring strips share that XCD's L2
    const int work = a.strips * a.segs, per = (work + 7) / 8;
    const int item = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (item >= work) return;   // whole workgroup, before any barrier
    const int y0 = (item / a.strips) * a.seg_rows;
    const int y1 = min(y0 + a.seg_rows, g.Ny);
    const int x0 = (item % a.strips) * a.tx;
    // every wavefront runs the same number of iterations (a multiple of P) with the
    // same barriers; iterations past a level's rows do nothing
    const int iters = ((y1 - y0) + 3 * (D - 1) + P - 1) / P * P;
    const int rbase = y0 - (D - 1);

    if (level == 1) {
        // the loads are unconditional (rows past the last are clamped to it), so the
        // compiler's wait counts let P rows of loads stay in flight
        const int n1 = (y1 - y0) + 2 * (D - 1);   // level-1 rows
        const int x = ((x0 - (D - 1) + lane) % g.Nx + g.Nx) % g.Nx;
        const unsigned bx = (unsigned)x * (unsigned)sizeof(T);
        const unsigned bxm = (unsigned)(x == 0 ? g.Nx - 1 : x - 1) * (unsigned)sizeof(T);
        const unsigned bxp = (unsigned)(x == g.Nx - 1 ? 0 : x + 1) * (unsigned)sizeof(T);
        // periodic row of level-1 iteration t; the rows lie within D - 1 of the
        // lattice, so the wrap is a few conditional additions (no integer division)
        auto row = [&](int t) {
            int r = rbase + min(t, n1 - 1);
            while (r < 0) r += g.Ny;
            while (r >= g.Ny) r -= g.Ny;
            return r;
        };
        T pre[P][Q];
#pragma unroll
        for (int j = 0; j < P; j++) gather_row(S, g, row(j), bxm, bx, bxp, pre[j]);
        for (int t0 = 0; t0 < iters; t0 += P) {
#pragma unroll
            for (int j = 0; j < P; j++) {
                const int t = t0 + j;
                if (t < n1) pull_collide<T, W, NF>(pre[j], g, a.Fx_lag, a.Fy_lag, a.force, a.Fx, a.Fy);
                if constexpr (R == 3) __syncthreads();
                if (t < n1) {
                    T *out = buf + lane;
#pragma unroll
                    for (int k = 0; k < Q; k++) out[Sh::slot(k, rbase + t, rbase) * LW] = pre[j][k];
                }
                gather_row(S, g, row(t + P), bxm, bx, bxp, pre[j]);
                __syncthreads();
            }
        }
    } else {
        stream_levels<T, W, D, LW, R, NF>(level, buf, S, Dst, a, lane, x0, y0, y1, iters,
                                      std::make_integer_sequence<int, D - 1>{});
    }
}

// ---- D steps per launch in one wavefront's registers --------------------------
// k_wave: each wavefront owns a strip of 64 columns (TX = 64 - 2(D-1) of them
// written) and a segment of rows, and walks down the segment computing, per
// row, every time level in program order: level 1 gathers row t from S in
// global memory, level l (2..D) row t - (l-1) from level l-1's last three rows
// held in registers -- a lane's x - 1 and x + 1 neighbours come from the lanes
// beside it (DPP wave shifts), rows from the registers -- and level D stores its
// row.  Nothing passes through LDS and no barrier is needed; every row of every
// level is computed (rows beyond a segment are the periodic lattice's own rows,
// needed as the levels' ring or computed and dropped), so all lanes run the same
// instructions.  Level l is valid on lanes l-1 .. 64-l, whose sources are valid
// lanes of level l-1.  Same arithmetic on the same values as D k_pull launches:
// bit-identical.
constexpr int kWaveWaves = 4;   // independent wavefronts per workgroup

// lane i receives lane i-1 (wave_shr:1) or lane i+1 (wave_shl:1) of v; the edge
// lane keeps its own value (it is outside the valid lanes of the next level)
template <int CTRL>
__device__ __forceinline__ float dpp_shift(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL,
                                                      0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, 0xf, 0xf,
                                               false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
constexpr int kDppShr1 = 0x138, kDppShl1 = 0x130;

// What a level keeps of its last rows for the next level: the newest row n's
// populations {0, 1, 3} and {2, 5, 6}, and row n-1's {2, 5, 6}.
template <typename T>
struct WaveRows {
    T c013[3], c256[3], p256[3];
};

// Gather of row q of level l from level l-1: row q+1 is `nw` (just computed), row q
// and row q-1 are what `h` kept.  Then `h` moves on to row q+1.
template <typename T>
__device__ __forceinline__ void wave_gather(WaveRows<T> &h, const T nw[Q], T f[Q])
{
    f[0] = h.c013[0];
    f[1] = dpp_shift<kDppShr1>(h.c013[1]);   // c = (+1, 0): from x-1
    f[3] = dpp_shift<kDppShl1>(h.c013[2]);   // c = (-1, 0): from x+1
    f[2] = h.p256[0];                        // c = (0, +1): from row q-1
    f[5] = dpp_shift<kDppShr1>(h.p256[1]);   // c = (+1, +1)
    f[6] = dpp_shift<kDppShl1>(h.p256[2]);   // c = (-1, +1)
    f[4] = nw[4];                            // c = (0, -1): from row q+1
    f[8] = dpp_shift<kDppShr1>(nw[8]);       // c = (+1, -1)
    f[7] = dpp_shift<kDppShl1>(nw[7]);       // c = (-1, -1)
    h.p256[0] = h.c256[0];
    h.p256[1] = h.c256[1];
    h.p256[2] = h.c256[2];
    h.c013[0] = nw[0];
    h.c013[1] = nw[1];
    h.c013[2] = nw[3];
    h.c256[0] = nw[2];
    h.c256[1] = nw[5];
    h.c256[2] = nw[6];
}
template <typename T, bool W, int D, int P, bool NF = false>
__global__ void __launch_bounds__(64 * kWaveWaves) k_wave(const T *__restrict__ S,
                                                         T *__restrict__ Dst, StreamArgs a)
{
    const Geo &g = a.g;
    const int lane = threadIdx.x % 64;
    // XCD-aware order: XCD i (workgroups with blockIdx % 8 == i) takes the i-th
    // eighth of the strip-major work list
    const int work = a.strips * a.segs;
    const int per = (int)gridDim.x / 8;
    const int item = (((int)blockIdx.x % 8) * per + (int)blockIdx.x / 8) * kWaveWaves +
                     __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    if (item >= work) return;   // whole wavefront (no barrier in this kernel)
    const int y0 = (item / a.strips) * a.seg_rows;
    const int y1 = min(y0 + a.seg_rows, g.Ny);
    const int x0 = (item % a.strips) * a.tx;
    const int x = ((x0 - (D - 1) + lane) % g.Nx + g.Nx) % g.Nx;
    const unsigned bx = (unsigned)x * (unsigned)sizeof(T);
    const unsigned bxm = (unsigned)(x == 0 ? g.Nx - 1 : x - 1) * (unsigned)sizeof(T);
    const unsigned bxp = (unsigned)(x == g.Nx - 1 ? 0 : x + 1) * (unsigned)sizeof(T);
    const int xo = x0 + lane - (D - 1);   // output column of this lane at level D
    const bool out_lane = lane >= D - 1 && lane < D - 1 + a.tx && xo < g.Nx;
    const int rbase = y0 - (D - 1);
    const int iters = (y1 - y0) + 2 * (D - 1);
    auto row = [&](int t) {   // periodic row of level-1 iteration t
        int r = rbase + t;
        while (r < 0) r += g.Ny;
        while (r >= g.Ny) r -= g.Ny;
        return r;
    };

    WaveRows<T> h[D - 1];
    T pre[P][Q];
#pragma unroll
    for (int j = 0; j < P; j++) gather_row(S, g, row(j), bxm, bx, bxp, pre[j]);
    for (int t0 = 0; t0 < iters; t0 += P) {
#pragma unroll
        for (int j = 0; j < P; j++) {
            const int t = t0 + j;
            T nw[Q], f[Q];
#pragma unroll
            for (int k = 0; k < Q; k++) nw[k] = pre[j][k];
            gather_row(S, g, row(t + P), bxm, bx, bxp, pre[j]);
            // level 1: row rbase + t
            pull_collide<T, W, NF>(nw, g, a.Fx_lag, a.Fy_lag, a.force, a.Fx, a.Fy);
            // levels 2 .. D: row rbase + t - (l-1)
#pragma unroll
            for (int l = 2; l <= D; l++) {
                wave_gather(h[l - 2], nw, f);
                pull_collide<T, W, NF>(f, g, a.Fx, a.Fy, a.force, a.Fx, a.Fy);
#pragma unroll
                for (int k = 0; k < Q; k++) nw[k] = f[k];
            }
            const int r = rbase + t - (D - 1);
            if (out_lane && r >= y0 && r < y1) {
                char *rowp = reinterpret_cast<char *>(Dst + (size_t)r * g.pitch);
                const unsigned bo = (unsigned)xo * (unsigned)sizeof(T);
#pragma unroll
                for (int k = 0; k < Q; k++)
                    store_nt(reinterpret_cast<T *>(rowp + (size_t)k * g.plane * sizeof(T) + bo),
                             nw[k]);
            }
        }
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
