// batch_kernels.h -- gfx950 kernels that advance many independent lattices in one
// persistent launch per time step.
//
// Storage of one lattice (one arena block): two population buffers of nine
// dense planes, node n = y*Nx + x, plane stride `plane` elements (a multiple of
// 64, so every plane starts 256-byte aligned).  The buffers hold POST-collision
// populations f* (gather form): a step reads f*_s from buf[s & 1], forms the
// pre-collision populations of the next step by pulling from the neighbours
// (streaming.c:21-42), applies the case's boundary closures, takes the moments,
// collides (collision.c:24-37) and writes f*_{s+1} into buf[(s + 1) & 1].
//
// Reference semantics kept exactly (src/core/main_lbm.c:31-61):
//   * solid sources give half-way bounce-back, f(x, k) = f*(x, opp k)
//     (bounce_back.c:43-46); solid nodes are never written.
//   * sources outside a non-periodic domain leave the entry stale
//     (streaming.c:36-37): it keeps f*_{s-1}(x, k), which is exactly what the
//     destination buffer holds before this step overwrites it.
//   * collision uses the moments left by the previous step, i.e. the moments of
//     the final (post-boundary) populations it collides; the first collision uses
//     the initial moment arrays (done by k_batch_init).
//   * boundary closures that read lagged moments (the regularized wall) get them
//     from the neighbour's previous step, kept in a small per-lattice array.
// The FP64 arithmetic is d2q9.h's (reference rounding); FP32-shifted storage uses
// step_kernels.h's moments_f32/collide_f32 and decodes to FP64 for the boundary
// closures, exactly as the single-lattice kernels do.
#pragma once
#include "step_kernels.h"
#include "tdibc_kernels.h"

namespace lbmx {

enum : unsigned {
    BF_PER_X = 1u << 0,
    BF_PER_Y = 1u << 1,
    BF_WALL_BOTTOM = 1u << 2,
    BF_WALL_TOP = 1u << 3,
    BF_WALL_LEFT = 1u << 4,
    BF_WALL_RIGHT = 1u << 5,
    BF_TOP_VELOCITY = 1u << 6,
    BF_CONVECTIVE = 1u << 7,
    BF_LINER = 1u << 8,
    BF_FORCE = 1u << 9,
    BF_MOMENTS = 1u << 10,
};

// Per-lattice descriptor, read with scalar loads (one lattice per tile).
struct BatchLat {
    void *buf[2];            // population buffers (T[9][plane])
    const double *om_col;    // relaxation rate per column
    double *out[3];          // rho, u, v at completion (BF_MOMENTS)
    double *saved[2];        // liner: lagged (rho, u, v) of row 1 over the opening, by parity
    double *state;           // liner: phi[2] psi[2] chi[2] rho_lag (8 doubles) per opening node
    double *probe_hist;      // (steps + 1) x nprobe
    unsigned *status;        // divergence guard bits (kGuard*), cleared by k_batch_init
    void *fbuf;              // frame buffer (T[9][fplane]): nodes within 3 of an edge (pairs)
    unsigned fplane;
    long long start;         // batch epoch at which this lattice takes its step 0
    unsigned plane;
    int Nx, Ny;
    unsigned flags;
    int gap_x0, gap_x1;
    int steps;
    int nprobe;
    int probe_n[4];
    double inv_nx;           // 1 / Nx: interior index -> row
    double rho0, Fx, Fy, U_top, V_top;
    TdibcParams tp;
};

struct BatchCtl {
    long long epoch;         // steps launched so far (k_batch_advance)
    int n_interior;          // interior tiles (tile table [0, n_interior))
    int n_frame;             // frame tiles (tile table [n_interior, n_interior + n_frame))
};

struct Tile {
    int lat;                 // descriptor slot
    int start;               // first node of the tile in its kind's numbering
};

// Initial state sources of k_batch_init.
struct BatchInit {
    int kind;                // LBMX_INIT_PULSE / _TGV / _FIELDS
    const double *prof;      // pulse: rho per column
    const double *tx;        // TGV: sin(kx x), cos(kx x), cos(2 kx x) per column (3 x Nx)
    const double *ty;        // TGV: sin(ky y), cos(ky y), cos(2 ky y) per row (3 x Ny)
    const double *fields;    // rho, u, v planes (3 x Nx*Ny)
    double u0, r;            // TGV amplitude and kx/ky
};

constexpr int kBatchThreads = 256;

// Pointers read from a descriptor are generic to the compiler; casting them to the
// global address space makes every access a global_load/global_store (flat ones
// would also count on the scalar-load counter, so each descriptor read would wait
// for all populations in flight).
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> gbl(T *p)
{
    return (gptr<T>)p;
}
template <typename T>
__device__ __forceinline__ gptr<const T> gbl(const T *p)
{
    return (gptr<const T>)p;
}

// Divergence guard on the moments every collision reads (the state after the
// previous step).  The reference drivers stop a run once max |u| > 0.2
// (cases/poiseuille/poiseuille.c:139-142, cases/couette/couette.c:145-148); here
// a lattice records every kind of violation it meets (the word is the union over
// its nodes and steps, so it does not depend on scheduling) and the batch keeps
// running the others, unchanged.
enum : unsigned { kGuardNaN = 1u, kGuardRho = 2u, kGuardSpeed = 4u };
constexpr double kGuardSpeedMax = 0.2;

// The density the moments return is floored at 1e-16 (compute_macroscopic_fields.c
// clamps it there), which the FP32 storage modes rebuild as rho0 + (floor - rho0),
// i.e. within one rounding of rho0 of the floor (0 or 1.1e-16 for rho0 = 1): so a
// non-positive population sum shows as a density at or below kGuardRhoMin.
constexpr double kGuardRhoMin = 1e-12;

__device__ __forceinline__ void guard_check(const BatchLat &L, double rho, double ux, double uy)
{
    const bool low = !(rho > kGuardRhoMin);
    const bool bad = low || !(fabs(ux) <= kGuardSpeedMax) || !(fabs(uy) <= kGuardSpeedMax);
    if (bad) {
        const bool nan = rho != rho || ux != ux || uy != uy;
        const unsigned code = nan ? kGuardNaN : (low ? kGuardRho : kGuardSpeed);
        // one atomic per wavefront: the union of its tripped lanes' codes
        const unsigned all = (__ballot(code == kGuardNaN) ? kGuardNaN : 0u) |
                             (__ballot(code == kGuardRho) ? kGuardRho : 0u) |
                             (__ballot(code == kGuardSpeed) ? kGuardSpeed : 0u);
        const unsigned long long lanes = __ballot(1);
        // skip the atomic once the word holds these bits (a stale read only costs
        // an atomic more), so a lattice gone bad everywhere does not contend for it
        if ((unsigned)__lane_id() == (unsigned)__builtin_ctzll(lanes) &&
            (*gbl((const unsigned *)L.status) & all) != all)
            __hip_atomic_fetch_or(gbl(L.status), all, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__device__ __forceinline__ bool b_solid(const BatchLat &L, int x, int y)
{
    const unsigned F = L.flags;
    return ((F & BF_WALL_BOTTOM) && y == 0 && (x < L.gap_x0 || x > L.gap_x1)) ||
           ((F & BF_WALL_TOP) && y == L.Ny - 1) || ((F & BF_WALL_LEFT) && x == 0) ||
           ((F & BF_WALL_RIGHT) && x == L.Nx - 1);
}

// stored value <-> population (FP32-shifted stores f - w*rho0, as pop_get/pop_put)
__device__ __forceinline__ double b_dec(double v, int, double) { return v; }
__device__ __forceinline__ double b_dec(float v, int k, double rho0)
{
    return (double)v + D2Q9::w[k] * rho0;
}
template <typename T>
__device__ __forceinline__ T b_enc(double v, int k, double rho0)
{
    if constexpr (sizeof(T) == 8)
        return v;
    else
        return (float)(v - D2Q9::w[k] * rho0);
}

__device__ __forceinline__ void b_store(gptr<double> p, double v)
{
    __builtin_nontemporal_store(v, p);
}
__device__ __forceinline__ void b_store(gptr<float> p, float v)
{
    __builtin_nontemporal_store(v, p);
}

// ---- where a step reads and writes populations -------------------------------------
// A view addresses population k of node (x, y) in some buffer of one lattice.
// FullV: a dense population buffer (nine planes of Nx*Ny nodes).  FrameV: the
// lattice's frame buffer, which holds only the nodes within 3 of an edge (rows
// 0-3 and Ny-4..Ny-1 whole, columns 0-3 and Nx-4..Nx-1 of the rows between; the
// whole lattice when Nx or Ny is below 9).  One step reads its sources through
// one view, the stale entries of a node's own slots through another and stores
// through a third (the single-step kernels: f*_s, and f*_{s+1}'s buffer for both
// the stale reads and the stores).
template <typename T>
struct FullV {
    gptr<T> p;
    unsigned P;
    int Nx;
    __device__ __forceinline__ T at(int k, int x, int y) const
    {
        return p[k * P + (unsigned)(y * Nx + x)];
    }
    __device__ __forceinline__ void put(int k, int x, int y, T v) const
    {
        b_store(&p[k * P + (unsigned)(y * Nx + x)], v);
    }
};

__host__ __device__ __forceinline__ bool frame_is_full(int Nx, int Ny)
{
    return Nx < 9 || Ny < 9;
}
__host__ __device__ __forceinline__ unsigned frame_nodes_of(int Nx, int Ny)
{
    return frame_is_full(Nx, Ny) ? (unsigned)(Nx * Ny) : (unsigned)(8 * Nx + 8 * (Ny - 8));
}
// Index of node (x, y) in the frame buffer (the node must lie within 3 of an edge).
__device__ __forceinline__ unsigned frame_index(int Nx, int Ny, int x, int y)
{
    if (frame_is_full(Nx, Ny) || y < 4) return (unsigned)(y * Nx + x);
    if (y >= Ny - 4) return (unsigned)((y - Ny + 8) * Nx + x);
    return (unsigned)(8 * Nx + (y - 4) * 8 + (x < 4 ? x : x - Nx + 8));
}

template <typename T>
struct FrameV {
    gptr<T> p;
    unsigned P;
    int Nx, Ny;
    __device__ __forceinline__ T at(int k, int x, int y) const
    {
        return p[k * P + frame_index(Nx, Ny, x, y)];
    }
    __device__ __forceinline__ void put(int k, int x, int y, T v) const
    {
        p[k * P + frame_index(Nx, Ny, x, y)] = v;
    }
};

template <typename T>
__device__ __forceinline__ FullV<T> full_view(const BatchLat &L, int b)
{
    return FullV<T>{gbl((T *)L.buf[b]), L.plane, L.Nx};
}
template <typename T>
__device__ __forceinline__ FrameV<T> frame_view(const BatchLat &L)
{
    return FrameV<T>{gbl((T *)L.fbuf), L.fplane, L.Nx, L.Ny};
}

// Pre-collision populations of node (x, y): pull from x - c_k with the wrap,
// bounce-back and stale rules above.  Stale entries are read from the node's own
// slots through O, which only this node's thread writes (callers gathering for
// another node never meet a stale entry; the host validates that).
template <typename T, class SV, class OV>
__device__ __forceinline__ void b_gather(const BatchLat &L, const SV &S, const OV &O, int x, int y,
                                         T f[Q])
{
    const int Nx = L.Nx, Ny = L.Ny;
    const bool px = L.flags & BF_PER_X, py = L.flags & BF_PER_Y;
    f[0] = S.at(0, x, y);
#pragma unroll
    for (int k = 1; k < Q; k++) {
        const int cx = D2Q9::cx[k], cy = D2Q9::cy[k];
        int xs = x - cx, ys = y - cy;
        bool oob = false;
        if (cx != 0) {
            if (xs < 0) {
                xs += Nx;
                oob = !px;
            } else if (xs >= Nx) {
                xs -= Nx;
                oob = !px;
            }
        }
        if (cy != 0) {
            if (ys < 0) {
                ys += Ny;
                oob = oob || !py;
            } else if (ys >= Ny) {
                ys -= Ny;
                oob = oob || !py;
            }
        }
        const bool bb = !oob && b_solid(L, xs, ys);
        f[k] = oob ? O.at(k, x, y) : (bb ? S.at(D2Q9::opp[k], x, y) : S.at(k, xs, ys));
    }
}

// Moments, collision and the node's outputs for node (x, y) whose pre-collision
// populations (after its boundary closures) are f, which then holds f*_{s+1}.
// own = false (a neighbour recomputed by another tile): the collision only.
template <typename T, bool W>
__device__ __forceinline__ void batch_update(const BatchLat &L, int s, int x, int y, int n, T f[Q],
                                             bool last, bool own = true)
{
    const unsigned F = L.flags;
    const double rho0 = L.rho0;
    // a lattice's force flag is uniform over its tiles: each branch below runs a
    // collision specialised for it (the force-free one without the Guo terms, 4.7 %
    // of the batched pass when no lattice has a force)
    const bool force = F & BF_FORCE;
    const double om = gbl(L.om_col)[x];
    const bool moments = own && last && (F & BF_MOMENTS);
    double rl, ul, vl;   // next step's lagged moments as the reference's arrays hold them
    if constexpr (sizeof(T) == 8) {
        double *fd = reinterpret_cast<double *>(f);
        const Moments m = moments_ref(fd, L.Fx, L.Fy);
        rl = m.rho;
        ul = m.ux;
        vl = m.uy;
        if (moments) {
            gbl(L.out[0])[n] = m.rho;
            gbl(L.out[1])[n] = m.ux;
            gbl(L.out[2])[n] = m.uy;
        }
        if (force)
            collide_ref(fd, m, om, true, L.Fx, L.Fy);
        else
            collide_nf(fd, m, om);   // collide_ref without a force, bit for bit
    } else if constexpr (W) {
        float *q = reinterpret_cast<float *>(f);
        const MomentsW m = moments_wide(q, rho0, L.Fx, L.Fy);
        rl = rho0 + m.dr;
        ul = m.ux;
        vl = m.uy;
        if (moments) {
            double fd[Q];
#pragma unroll
            for (int k = 0; k < Q; k++) fd[k] = b_dec(q[k], k, rho0);
            const Moments md = moments_ref(fd, L.Fx, L.Fy);
            gbl(L.out[0])[n] = md.rho;
            gbl(L.out[1])[n] = md.ux;
            gbl(L.out[2])[n] = md.uy;
        }
        if (force)
            collide_wide(q, m, rho0, om, true, L.Fx, L.Fy);
        else
            collide_wide(q, m, rho0, om, false, 0.0, 0.0);
    } else {
        float *q = reinterpret_cast<float *>(f);
        const float r0 = (float)rho0;
        const MomentsF m = moments_f32(q, r0, (float)L.Fx, (float)L.Fy);
        rl = (double)m.dr + rho0;
        ul = (double)m.ux;
        vl = (double)m.uy;
        if (moments) {
            double fd[Q];
#pragma unroll
            for (int k = 0; k < Q; k++) fd[k] = b_dec(q[k], k, rho0);
            const Moments md = moments_ref(fd, L.Fx, L.Fy);
            gbl(L.out[0])[n] = md.rho;
            gbl(L.out[1])[n] = md.ux;
            gbl(L.out[2])[n] = md.uy;
        }
        if (force)
            collide_f32(q, m, r0, (float)om, true, (float)L.Fx, (float)L.Fy);
        else
            collide_f32(q, m, r0, (float)om, false, 0.0f, 0.0f);
    }
    if (!own) return;

    guard_check(L, rl, ul, vl);

    if (F & BF_LINER) {
        if (y == 0 && x >= L.gap_x0 && x <= L.gap_x1) gbl(L.state)[8 * (x - L.gap_x0) + 6] = rl;
        if (y == 1 && x >= L.gap_x0 && x <= L.gap_x1) {
            gptr<double> sv = gbl(L.saved[(s + 1) & 1]) + 3 * (x - L.gap_x0);
            sv[0] = rl;
            sv[1] = ul;
            sv[2] = vl;
        }
    }
    for (int i = 0; i < L.nprobe; i++)
        if (n == L.probe_n[i]) gbl(L.probe_hist)[(long long)(s + 1) * L.nprobe + i] = rl;
}

// Node (x, y) of rows 1 .. Ny-2.  Away from the edge columns (2 <= x <= Nx-3,
// 2 <= y <= Ny-3) every source x - c_k is an in-domain fluid node (walls occupy
// only the outermost rows and columns): a plain gather.  Elsewhere the general
// gather applies wrap, bounce-back and stale entries, and a convective end node
// takes its inner neighbour's populations; no other closure touches these rows.
// Every lane stores all nine slots (a solid node stores zeros: its slots are
// never read), so a wavefront writes whole cache lines.
// The general gather of a row node: wrap, bounce-back, stale entries, and a
// convective end node taking its inner neighbour's populations.  A solid node
// gets zeros (and NaN completion moments) and returns false.
template <typename T, class SV, class OV>
__device__ __forceinline__ bool edge_gather(const BatchLat &L, const SV &S, const OV &O, int x,
                                            int y, bool last, bool own, T f[Q])
{
    const int Nx = L.Nx;
    const unsigned F = L.flags;
    if (b_solid(L, x, y)) {
        if (own && last && (F & BF_MOMENTS)) {
            const int n = y * Nx + x;
            gbl(L.out[0])[n] = NAN;
            gbl(L.out[1])[n] = NAN;
            gbl(L.out[2])[n] = NAN;
        }
#pragma unroll
        for (int k = 0; k < Q; k++) f[k] = (T)0;
        return false;
    }
    int gx = x;
    if ((F & BF_CONVECTIVE) && (x == 0 || x == Nx - 1)) gx = (x == 0) ? 1 : Nx - 2;
    b_gather<T>(L, S, O, gx, y, f);
    return true;
}

template <typename T, bool W, class SV, class OV>
__device__ __forceinline__ void row_node_update(const BatchLat &L, const SV &S, const OV &O, int s,
                                                int x, int y, bool last, bool own, T f[Q])
{
    const int Nx = L.Nx, Ny = L.Ny;
    bool fluid = true;
    if ((unsigned)(x - 2) < (unsigned)(Nx - 4) && (unsigned)(y - 2) < (unsigned)(Ny - 4)) {
#pragma unroll
        for (int k = 0; k < Q; k++) f[k] = S.at(k, x - D2Q9::cx[k], y - D2Q9::cy[k]);
    } else {
        fluid = edge_gather<T>(L, S, O, x, y, last, own, f);
    }
    if (fluid) batch_update<T, W>(L, s, x, y, y * Nx + x, f, last, own);
}

template <typename T, bool W, class SV, class OV, class DV>
__device__ __forceinline__ void batch_row_node(const BatchLat &L, const SV &S, const OV &O,
                                               const DV &D, int s, int x, int y, bool last)
{
    T f[Q];
    row_node_update<T, W>(L, S, O, s, x, y, last, true, f);
#pragma unroll
    for (int k = 0; k < Q; k++) D.put(k, x, y, f[k]);
}

// Node of row 0 or Ny-1: wrap, walls, stale entries and the case's boundary
// closures (moving lid, impedance wall, convective corners).
template <typename T, bool W, class SV, class OV, class DV>
__device__ __forceinline__ void batch_frame(const BatchLat &L, const SV &S, const OV &O,
                                            const DV &D, int s, int x, int y)
{
    const int Nx = L.Nx, Ny = L.Ny;
    const int n = y * Nx + x;
    const unsigned F = L.flags;
    const bool last = (s + 1 == L.steps);
    if (b_solid(L, x, y)) {
        if (last && (F & BF_MOMENTS)) {
            gbl(L.out[0])[n] = NAN;
            gbl(L.out[1])[n] = NAN;
            gbl(L.out[2])[n] = NAN;
        }
        return;
    }
    const double rho0 = L.rho0;

    // convective_{inlet,outlet}: the node takes the post-streaming populations of
    // its inner neighbour (convective_inlet.c:20-27); the impedance wall node
    // starts from those of the node above it (regularized closure input).
    int gx = x, gy = y;
    const bool conv = (F & BF_CONVECTIVE) && (x == 0 || x == Nx - 1);
    if (conv) gx = (x == 0) ? 1 : Nx - 2;
    const bool liner = (F & BF_LINER) && y == 0 && x >= L.gap_x0 && x <= L.gap_x1;
    if (liner) gy = 1;
    T f[Q];
    b_gather<T>(L, S, O, gx, gy, f);

    if (liner) {
        // cases/nit/nit.c:13-47 per node, then velocity_bottom_regularized.c:22-70
        const int e = x - L.gap_x0;
        gptr<double> st = gbl(L.state) + 8 * e;
        TdibcState ts;
        ts.phi[0] = st[0];
        ts.phi[1] = st[1];
        ts.psi[0] = st[2];
        ts.psi[1] = st[3];
        ts.chi[0] = st[4];
        ts.chi[1] = st[5];
        const TdibcParams &P = L.tp;
        const double p = P.cs * P.cs * (st[6] - P.rho0) * P.cp;
        const double v = ade_velocity(ts, P, p);
        st[0] = ts.phi[0];
        st[1] = ts.phi[1];
        st[2] = ts.psi[0];
        st[3] = ts.psi[1];
        st[4] = ts.chi[0];
        st[5] = ts.chi[1];
        const double vbc = P.gain * (v / P.cu);
        const gptr<const double> lg = gbl((const double *)L.saved[s & 1]) + 3 * e;
        double fi[Q], fb[Q];
#pragma unroll
        for (int k = 0; k < Q; k++) fi[k] = b_dec(f[k], k, rho0);
        regularized_node(fi, lg[0], lg[1], lg[2], lg[0], 0.0, vbc, fb);
#pragma unroll
        for (int k = 0; k < Q; k++) f[k] = b_enc<T>(fb[k], k, rho0);
    }
    if ((F & BF_TOP_VELOCITY) && y == Ny - 1) {
        double fd[Q];
#pragma unroll
        for (int k = 0; k < Q; k++) fd[k] = b_dec(f[k], k, rho0);
        zou_he_top_node(fd, L.U_top, L.V_top);
        f[4] = b_enc<T>(fd[4], 4, rho0);
        f[7] = b_enc<T>(fd[7], 7, rho0);
        f[8] = b_enc<T>(fd[8], 8, rho0);
    }
    batch_update<T, W>(L, s, x, y, n, f, last);
#pragma unroll
    for (int k = 0; k < Q; k++) D.put(k, x, y, f[k]);
}

// Node counts of the two tile kinds of a lattice (Nx, Ny >= 4): row tiles run over
// whole rows 1 .. Ny-2, edge tiles over rows 0 and Ny-1.
__host__ __device__ __forceinline__ int interior_nodes(int Nx, int Ny)
{
    return Nx * (Ny - 2);
}
__host__ __device__ __forceinline__ int frame_nodes(int Nx, int Ny)
{
    return 2 * Nx;
}
constexpr int kFrameTile = 256;
constexpr int kFrameThreads = 256;

// Tiles of workgroup b out of G for n tiles, as the k-th tile of b's sequence.
// mode 0: interleaved (tile t on workgroup t mod G, so the chip sweeps the table
// as one front, as a hardware-dispatched grid would); 1: one contiguous band per
// XCD (blocks b with equal b % 8 run on one XCD); 2: one contiguous range per
// workgroup; 3 and 4: interleaved bands of 4 / 16 consecutive tiles, each band on
// one XCD (band c on XCD c % 8, its tiles dealt round-robin to that XCD's
// workgroups), so a tile's neighbouring rows are mostly read through the same L2
// while the chip still sweeps the table as one front; 5: one tile per workgroup
// (the grid covers the table, rounded up: workgroups past the end do nothing).
// Mode 5 measured fastest on every mix tried (one 8192^2 lattice, 1,024 mixed
// lattices, 1,024 lattices of 256^2: 13-19 % more bandwidth than the persistent
// modes) and is the default; 0-4 remain for comparison.
struct TileSeq {
    int mode, b, G, n, t0, t1, B;
    __device__ __forceinline__ int at(int k) const
    {
        if (mode == 5) return k == 0 ? b : n;
        if (mode == 0) return b + k * G;
        if (mode == 1 || mode == 2) return t0 + k < t1 ? t0 + k : n;
        const int G8 = G / 8, x = b % 8, q = b / 8 + k * G8;
        return (x + 8 * (q / B)) * B + q % B;
    }
};
__device__ __forceinline__ TileSeq tile_seq(int n, int mode)
{
    TileSeq s;
    s.mode = mode;
    s.b = blockIdx.x;
    s.G = gridDim.x;
    s.n = n;
    s.B = mode == 3 ? 4 : 16;
    s.t0 = s.t1 = 0;
    if (mode == 1 || mode == 2) {
        const int chunk = mode == 1 ? (s.b % 8) * (s.G / 8) + s.b / 8 : s.b;
        s.t0 = (int)((long long)n * chunk / s.G);
        s.t1 = (int)((long long)n * (chunk + 1) / s.G);
    }
    return s;
}

// One time step of every lattice whose own step lies in [0, steps): rows 1 .. Ny-2,
// `tile` consecutive nodes of those rows per tile, tiles dealt by `mode`.
template <typename T, bool W>
__global__ void __launch_bounds__(kBatchThreads) k_batch_interior(
    const BatchLat *__restrict__ lats, const Tile *__restrict__ tiles,
    const BatchCtl *__restrict__ ctl, int local, int tile, int mode)
{
    const long long epoch = ctl->epoch + local;
    const int n = ctl->n_interior;
    const TileSeq Q = tile_seq(n, mode);
    for (int k = 0;; k++) {
        const int t = Q.at(k);   // increasing in k
        if (t >= n) break;
        const Tile tl = tiles[t];
        const BatchLat &L = lats[tl.lat];
        const long long s = epoch - L.start;
        if (s < 0 || s >= L.steps) continue;
        const int Nx = L.Nx;
        const int end = min(tl.start + tile, interior_nodes(Nx, L.Ny));
        const FullV<T> S = full_view<T>(L, (int)(s & 1)), D = full_view<T>(L, (int)((s + 1) & 1));
        const bool last = (s + 1 == L.steps);
        for (int m = tl.start + (int)threadIdx.x; m < end; m += (int)blockDim.x) {
            // exact for m < 2^26: (m + 1/2)/Nx is >= 1/(2 Nx) away from an integer
            const int r = (int)(((double)m + 0.5) * L.inv_nx);
            batch_row_node<T, W>(L, S, D, D, (int)s, m - r * Nx, 1 + r, last);
        }
    }
}

// The same step for rows 0 and Ny-1.  Independent of k_batch_interior within a
// step: the two write disjoint nodes and read only the previous step's buffer
// (plus a node's own stale slots).
template <typename T, bool W>
__global__ void __launch_bounds__(kFrameThreads) k_batch_frame(
    const BatchLat *__restrict__ lats, const Tile *__restrict__ tiles,
    const BatchCtl *__restrict__ ctl, int local)
{
    const long long epoch = ctl->epoch + local;
    const int base = ctl->n_interior;
    const int n = ctl->n_frame;
    const TileSeq Q = tile_seq(n, 0);
    for (int k = 0;; k++) {
        const int t = Q.at(k);
        if (t >= n) break;
        const Tile tl = tiles[base + t];
        const BatchLat &L = lats[tl.lat];
        const long long s = epoch - L.start;
        if (s < 0 || s >= L.steps) continue;
        const int Nx = L.Nx;
        const int end = min(tl.start + kFrameTile, frame_nodes(Nx, L.Ny));
        for (int f = tl.start + (int)threadIdx.x; f < end; f += (int)blockDim.x) {
            const int top = f >= Nx;
            const FullV<T> S = full_view<T>(L, (int)(s & 1)), D = full_view<T>(L, (int)((s + 1) & 1));
            batch_frame<T, W>(L, S, D, D, (int)s, f - top * Nx, top ? L.Ny - 1 : 0);
        }
    }
}

// ---- two steps per pass (pair mode) -------------------------------------------------
// A lattice's nodes split into the core -- 3 <= x <= Nx-4 and 3 <= y <= Ny-4, where
// a node and every source of its gather obey the plain rule -- and the band, the
// nodes within 2 of an edge, where every boundary rule and closure lives.  A wide
// core (wide_core) also takes the side columns of rows 3 .. Ny-4, whose nodes
// then run the general gather inside the core tiles, and leaves the band rows
// 0..2 and Ny-3..Ny-1.  With the state f*_s in buf[X], X = (s / 2) & 1 (pairs
// start at even s), and Y = 1 - X:
//   k_pair_band<0>  step s+1 of the band nodes: sources buf[X], stale entries
//                   (f*_{s-1}) from the frame buffer's own slots, stores to the
//                   frame buffer;
//   k_pair_core     step s+1 of a core tile and a one-node ring in LDS (each
//                   neighbour recomputes the ring), the tile's nodes next to the
//                   band also to the frame buffer, then step s+2 of the tile,
//                   stored to buf[Y];
//   k_pair_band<1>  step s+2 of the band nodes: sources from the frame buffer
//                   (band nodes and the core's outer ring), stale entries (f*_s)
//                   from buf[X]'s own slots, stores to buf[Y].
// The frame buffer then again holds f*_{s+1} at the band nodes, the stale values
// of the next pair.  Each node of each step is computed once by its owner with the
// arithmetic of the single-step kernels, so results are bit-identical to them.
// A lattice whose last step is s+1 takes one step: the first band phase and the
// core kernel store f*_{s+1} to buf[Y] and the second band phase skips it.
constexpr int kPairTX = 64;
// core tile rows and workgroup size by stored type: FP32 storage 64 x 16 tiles on
// 512 lanes (42.8 KB of LDS, six waves per SIMD; 64 x 8 on 256 lanes recomputes
// 29 % of the nodes instead of 16 %), FP64 64 x 8 on 256 lanes (47.5 KB)
template <typename T> struct PairShape {
    static constexpr int TY = sizeof(T) == 8 ? 8 : 16;
    static constexpr int threads = sizeof(T) == 8 ? 256 : 512;
};
__host__ __device__ __forceinline__ int pair_ty(int esize) { return esize == 8 ? 8 : 16; }
constexpr int kBandTile = 256;

__host__ __device__ __forceinline__ int core_w(int Nx) { return Nx > 6 ? Nx - 6 : 0; }
__host__ __device__ __forceinline__ int core_h(int Ny) { return Ny > 6 ? Ny - 6 : 0; }
// A lattice whose side columns meet no stale entry in rows 3 .. Ny-4 -- each side
// wraps, is a wall or is a convective end -- has a wide core: its core tiles span
// every column (the side nodes take the general gather, from LDS in the second
// step), and the band is rows 0..2 and Ny-3..Ny-1 only.  The three-node side
// columns of a narrow core cost a cache line per plane and row each way.
// A convective end node gathers around its inner neighbour, so the tile holding
// column Nx-1 must also hold Nx-3 in its ring: not so when that tile is one
// column wide (Nx = 64 t + 1), which keeps a narrow core.
__host__ __device__ __forceinline__ bool wide_core(unsigned F, int Nx, int Ny)
{
    if (frame_is_full(Nx, Ny)) return false;
    if ((F & BF_CONVECTIVE) && Nx % kPairTX == 1) return false;
    return (F & (BF_PER_X | BF_WALL_LEFT | BF_CONVECTIVE)) &&
           (F & (BF_PER_X | BF_WALL_RIGHT | BF_CONVECTIVE));
}
__host__ __device__ __forceinline__ int band_nodes(int Nx, int Ny, bool wide)
{
    return wide ? 6 * Nx : Nx * Ny - core_w(Nx) * core_h(Ny);
}
// Core tiles cover columns [64 t, 64 t + 64) (cut to the core, or to the lattice
// when it is wide), so that a tile's rows start on a 128-byte line when Nx is a
// multiple of 32, and rows [3 + ty j, 3 + ty (j + 1)).
__host__ __device__ __forceinline__ int core_tiles_x(int Nx, bool wide)
{
    if (wide) return (Nx + kPairTX - 1) / kPairTX;
    return core_w(Nx) ? (Nx - 4 + kPairTX - 1) / kPairTX : 0;
}
__host__ __device__ __forceinline__ int core_tiles(int Nx, int Ny, int ty, bool wide)
{
    return core_tiles_x(Nx, wide) * ((core_h(Ny) + ty - 1) / ty);
}

// Workgroup -> item of a table of n items: XCD i (workgroups with blockIdx % 8 ==
// i) takes the i-th eighth of the table in order, so neighbouring tiles -- whose
// halo rows overlap -- are read through the same L2.  The grid is a multiple of 8.
__device__ __forceinline__ int xcd_item()
{
    const int per = (int)gridDim.x / 8;
    return ((int)blockIdx.x % 8) * per + (int)blockIdx.x / 8;
}

// Band node m (row-major over the band) -> (x, y).
__device__ __forceinline__ void band_node(int Nx, int Ny, bool wide, int m, int &x, int &y)
{
    const int ch = core_h(Ny);
    if (!wide && (core_w(Nx) == 0 || ch == 0)) {
        y = m / Nx;
        x = m - y * Nx;
        return;
    }
    const int side = wide ? 0 : 6 * ch;   // side-column nodes of a narrow core
    if (m < 3 * Nx) {
        y = m / Nx;
        x = m - y * Nx;
    } else if (m < 3 * Nx + side) {
        const int q = m - 3 * Nx, r = q / 6, c = q - 6 * r;
        y = 3 + r;
        x = c < 3 ? c : Nx - 6 + c;
    } else {
        const int q = m - 3 * Nx - side, r = q / Nx;
        y = Ny - 3 + r;
        x = q - r * Nx;
    }
}

// A core tile's step-s+1 values in LDS: column lx = x - x0 + 1 (the ring's columns
// of a periodic lattice hold the wrapped nodes), row ly = y - y0 + 1.
template <typename T, int RX, int RN>
struct TileV {
    const T *b;
    int x0, y0, Nx;
    __device__ __forceinline__ T at(int k, int x, int y) const
    {
        int lx = x - x0 + 1;
        if (lx < 0)
            lx += Nx;
        else if (lx >= RX)
            lx -= Nx;
        return b[k * RN + (y - y0 + 1) * RX + lx];
    }
};

// The gather of a node in rows 2 .. Ny-3 of a wide core: b_gather's rules where no
// source can be out of range or in a solid row -- the x axis wraps, or the side
// is a wall (bounce-back from the node's own slots), or a convective end node
// gathers around its inner neighbour; on a plain node (2 <= x <= Nx-3) it is the
// plain gather.  A wall node gets zeros (and NaN completion moments) and returns
// false.
template <typename T, class SV>
__device__ __forceinline__ bool side_gather(const BatchLat &L, const SV &S, int x, int y, bool last,
                                            bool own, T f[Q])
{
    const int Nx = L.Nx;
    const unsigned F = L.flags;
    const bool wl = F & BF_WALL_LEFT, wr = F & BF_WALL_RIGHT;
    if ((wl && x == 0) || (wr && x == Nx - 1)) {
        if (own && last && (F & BF_MOMENTS)) {
            const int n = y * Nx + x;
            gbl(L.out[0])[n] = NAN;
            gbl(L.out[1])[n] = NAN;
            gbl(L.out[2])[n] = NAN;
        }
#pragma unroll
        for (int k = 0; k < Q; k++) f[k] = (T)0;
        return false;
    }
    int gx = x;
    if ((F & BF_CONVECTIVE) && (x == 0 || x == Nx - 1)) gx = (x == 0) ? 1 : Nx - 2;
    f[0] = S.at(0, gx, y);
#pragma unroll
    for (int k = 1; k < Q; k++) {
        int xs = gx - D2Q9::cx[k];
        if (xs < 0)
            xs += Nx;
        else if (xs >= Nx)
            xs -= Nx;
        // bounce-back reads the node's own opposite slot: one load either way
        const bool bb = (wl && xs == 0) || (wr && xs == Nx - 1);
        f[k] = S.at(bb ? D2Q9::opp[k] : k, bb ? gx : xs, bb ? y : y - D2Q9::cy[k]);
    }
    return true;
}

// Six waves per SIMD is what the tile's LDS allows (three FP32 workgroups per CU);
// the hint keeps the side-node gathers from costing a wave (at most 80 VGPRs).
template <typename T, bool W, int TY = PairShape<T>::TY, int NT = PairShape<T>::threads>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) k_pair_core(const BatchLat *__restrict__ lats,
                                                           const Tile *__restrict__ tiles,
                                                           const BatchCtl *__restrict__ ctl,
                                                           int local)
{
    constexpr int RX = kPairTX + 2, RN = RX * (TY + 2);
    __shared__ T buf[Q * RN];
    const long long epoch = ctl->epoch + local;
    const int item = xcd_item();
    if (item >= ctl->n_interior) return;
    const Tile tl = tiles[item];
    const BatchLat &L = lats[tl.lat];
    const long long s = epoch - L.start;
    if (s < 0 || s >= L.steps) return;
    const bool single = (s + 1 == L.steps);
    const int Nx = L.Nx, Ny = L.Ny;
    const bool wide = wide_core(L.flags, Nx, Ny);
    const bool px = L.flags & BF_PER_X;
    const int ntx = core_tiles_x(Nx, wide);
    const int x0 = (tl.start % ntx) * kPairTX, y0 = 3 + (tl.start / ntx) * TY;
    // the tile's own nodes: columns [xa, xb], rows [y0, min(y0 + TY, Ny - 3))
    const int xa = wide ? x0 : max(x0, 3);
    const int xb = wide ? min(x0 + kPairTX, Nx) - 1 : min(x0 + kPairTX - 1, Nx - 4);
    const int X = (int)((s >> 1) & 1);
    const FullV<T> S = full_view<T>(L, X), D = full_view<T>(L, X ^ 1);
    const FrameV<T> R = frame_view<T>(L);
    // A tile whose ring reaches a side node (x < 2 or x > Nx-3: the ring spans
    // x0-1 .. x0+64) lists its side nodes first and its plain nodes after them;
    // the first whole wavefronts take the side rule (which on a plain node is the
    // plain gather), the rest the plain one, so no wavefront runs both branches
    // and the item count, hence the passes per thread, is an interior tile's.
    const bool edge = wide && (x0 < 2 || x0 + kPairTX >= Nx - 2);

    // step s+1 of the node in ring column xu, LDS row ly
    auto step1 = [&](int xu, int ly, bool side_rule) {
        const int y = y0 - 1 + ly;
        if (y > Ny - 3) return;
        int x = xu;
        if (x < 0 || x >= Nx) x += x < 0 ? Nx : -Nx;   // (only on a periodic lattice)
        const bool own = xu >= xa && xu <= xb && ly >= 1 && ly <= TY && y <= Ny - 4;
        if (single && !own) return;
        const int r = ly * RX + (xu - x0 + 1);
        T p[Q];
        bool fluid = true;
        if (!side_rule) {
#pragma unroll
            for (int k = 0; k < Q; k++) p[k] = S.at(k, x - D2Q9::cx[k], y - D2Q9::cy[k]);
        } else {
            fluid = side_gather<T>(L, S, x, y, single, own, p);
        }
        if (fluid) batch_update<T, W>(L, (int)s, x, y, y * Nx + x, p, single, own);
        if (single) {
#pragma unroll
            for (int k = 0; k < Q; k++) D.put(k, x, y, p[k]);
            return;
        }
        if (own && (y == 3 || y == Ny - 4 || (!wide && (x == 3 || x == Nx - 4)))) {
#pragma unroll
            for (int k = 0; k < Q; k++) R.put(k, x, y, p[k]);
        }
#pragma unroll
        for (int k = 0; k < Q; k++) buf[k * RN + r] = p[k];
    };

    // step s+1: the tile and its one-node ring, columns xa-1 .. xb+1 (wrapped on a
    // periodic lattice; a narrow core's ring is plain, 2 <= x <= Nx-3)
    if (!edge) {
        for (int r = threadIdx.x; r < RN; r += NT) {
            const int lx = r % RX, xu = x0 - 1 + lx;
            if (xu >= xa - 1 && xu <= xb + 1) step1(xu, r / RX, false);
        }
    } else {
        // side columns [la, 1] and [Nx-2, rb], plain columns [pa, pb] between
        const int la = px ? -1 : 0, rb = min(px ? Nx : Nx - 1, xb + 1);
        const int nl = x0 == 0 ? 2 - la : 0, nr = xb + 1 >= Nx - 2 ? rb - (Nx - 2) + 1 : 0;
        const int pa = max(x0 - 1, 2), pb = min(xb + 1, Nx - 3), pw = max(pb - pa + 1, 0);
        const int ns = (nl + nr) * (TY + 2), n = ns + pw * (TY + 2);
        const int nsr = min((ns + 63) & ~63, n);   // whole wavefronts take the side rule
        for (int i = threadIdx.x; i < n; i += NT) {
            int xu, ly;
            if (i < ns) {
                const int c = i % (nl + nr);
                xu = c < nl ? la + c : Nx - 2 + (c - nl);
                ly = i / (nl + nr);
            } else {
                const int j = i - ns;
                xu = pa + j % pw;
                ly = j / pw;
            }
            step1(xu, ly, i < nsr);
        }
    }
    if (single) return;   // the whole workgroup: the lattice's last step was this one
    __syncthreads();

    // step s+2 of tile node x, row ty, from LDS
    const bool last = (s + 2 == L.steps);
    const TileV<T, RX, RN> V{buf, x0, y0, Nx};
    auto step2 = [&](int x, int ty, bool side_rule) {
        const int y = y0 + ty, tx = x - x0;
        if (y > Ny - 4) return;
        T p[Q];
        bool fluid = true;
        if (!side_rule) {
#pragma unroll
            for (int k = 0; k < Q; k++)
                p[k] = buf[k * RN + (ty + 1 - D2Q9::cy[k]) * RX + (tx + 1 - D2Q9::cx[k])];
        } else {
            fluid = side_gather<T>(L, V, x, y, last, true, p);
        }
        if (fluid) batch_update<T, W>(L, (int)s + 1, x, y, y * Nx + x, p, last);
#pragma unroll
        for (int k = 0; k < Q; k++) D.put(k, x, y, p[k]);
    };
    if (!edge) {
        for (int q = threadIdx.x; q < kPairTX * TY; q += NT) {
            const int x = x0 + q % kPairTX;
            if (x >= xa && x <= xb) step2(x, q / kPairTX, false);
        }
    } else {
        // side columns [xa, 1] and [max(Nx-2, xa), xb], plain columns between
        const int nl = xa <= 1 ? 2 - xa : 0;
        const int ra = max(Nx - 2, xa), nr = xb >= ra ? xb - ra + 1 : 0;
        const int qa = max(xa, 2), qw = max(min(xb, Nx - 3) - qa + 1, 0);
        const int ns = (nl + nr) * TY, n = ns + qw * TY;
        const int nsr = min((ns + 63) & ~63, n);
        for (int i = threadIdx.x; i < n; i += NT) {
            int x, ty;
            if (i < ns) {
                const int c = i % (nl + nr);
                x = c < nl ? xa + c : ra + (c - nl);
                ty = i / (nl + nr);
            } else {
                const int j = i - ns;
                x = qa + j % qw;
                ty = j / qw;
            }
            step2(x, ty, i < nsr);
        }
    }
}

// Band node (x, y) with the views of one phase (frame rows, or rows 1 .. Ny-2).
template <typename T, bool W, class SV, class OV, class DV>
__device__ __forceinline__ void band_step(const BatchLat &L, const SV &S, const OV &O, const DV &D,
                                          int s, int x, int y)
{
    if (y == 0 || y == L.Ny - 1)
        batch_frame<T, W>(L, S, O, D, s, x, y);
    else
        batch_row_node<T, W>(L, S, O, D, s, x, y, s + 1 == L.steps);
}

template <typename T, bool W, int PHASE>
__global__ void __launch_bounds__(kBandTile) k_pair_band(const BatchLat *__restrict__ lats,
                                                        const Tile *__restrict__ tiles,
                                                        const BatchCtl *__restrict__ ctl,
                                                        int local)
{
    const long long epoch = ctl->epoch + local;
    const int item = xcd_item();
    if (item >= ctl->n_frame) return;
    const Tile tl = tiles[ctl->n_interior + item];
    const BatchLat &L = lats[tl.lat];
    const long long s = epoch - L.start;
    if (s < 0 || s >= L.steps) return;
    const bool single = (s + 1 == L.steps);
    if (PHASE == 1 && single) return;
    const int m = tl.start + (int)threadIdx.x;
    const bool wide = wide_core(L.flags, L.Nx, L.Ny);
    if (m >= band_nodes(L.Nx, L.Ny, wide)) return;
    int x, y;
    band_node(L.Nx, L.Ny, wide, m, x, y);
    const int X = (int)((s >> 1) & 1);
    const FullV<T> A = full_view<T>(L, X), B = full_view<T>(L, X ^ 1);
    const FrameV<T> R = frame_view<T>(L);
    if (PHASE == 0) {
        if (single)
            band_step<T, W>(L, A, R, B, (int)s, x, y);
        else
            band_step<T, W>(L, A, R, R, (int)s, x, y);
    } else {
        band_step<T, W>(L, R, A, B, (int)s + 1, x, y);
    }
}

static __global__ void k_batch_advance(BatchCtl *ctl, int n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) ctl->epoch += n;
}

// Initial state of one lattice: buf[1] = feq(M0) (what the reference's f_new holds
// before the first step), buf[0] = collide(feq(M0)) with the lagged moments M0,
// probe row 0, the impedance state and the first lagged row-1 moments.
template <typename T, bool W>
__global__ void __launch_bounds__(256) k_batch_init(const BatchLat *__restrict__ lats, int slot,
                                                    BatchInit in)
{
    const BatchLat &L = lats[slot];
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= L.Nx) return;
    const int n = y * L.Nx + x;
    if (n == 0) *L.status = 0u;
    double rho, ux, uy;
    if (in.kind == 0) {
        rho = in.prof[x];
        ux = 0.0;
        uy = 0.0;
    } else if (in.kind == 1) {
        // lbmx_init_taylor_green's formulas on host-computed sin/cos tables
        const double sx = in.tx[x], cxx = in.tx[L.Nx + x], c2x = in.tx[2 * L.Nx + x];
        const double sy = in.ty[y], cyy = in.ty[L.Ny + y], c2y = in.ty[2 * L.Ny + y];
        ux = in.u0 * sx * cyy;
        uy = -in.u0 * in.r * cxx * sy;
        rho = 1.0 - 0.75 * in.u0 * in.u0 * (c2x + in.r * in.r * c2y);
    } else {
        const size_t N = (size_t)L.Nx * L.Ny;
        rho = in.fields[n];
        ux = in.fields[N + n];
        uy = in.fields[2 * N + n];
    }
    const double rho0 = L.rho0;
    const double usq = ux * ux + uy * uy;
    T *A = (T *)L.buf[0], *B = (T *)L.buf[1];
    const unsigned P = L.plane;
    T q[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) {
        const double fe = feq_ref(D2Q9::w[k], (double)D2Q9::cx[k], (double)D2Q9::cy[k], rho, ux,
                                  uy, usq);
        q[k] = b_enc<T>(fe, k, rho0);
        B[k * P + n] = q[k];
    }
    // the frame buffer's nodes: the same stale values for pair mode's first step
    if (L.fbuf && (x < 4 || y < 4 || x >= L.Nx - 4 || y >= L.Ny - 4 ||
                   frame_is_full(L.Nx, L.Ny))) {
        T *Rb = (T *)L.fbuf;
        const unsigned ri = frame_index(L.Nx, L.Ny, x, y);
#pragma unroll
        for (int k = 0; k < Q; k++) Rb[k * L.fplane + ri] = q[k];
    }
    const bool solid = b_solid(L, x, y);
    if (!solid) {
        const bool force = L.flags & BF_FORCE;
        const double om = L.om_col[x];
        if constexpr (sizeof(T) == 8) {
            const Moments m{rho, ux, uy};
            if (force)
                collide_ref(reinterpret_cast<double *>(q), m, om, true, L.Fx, L.Fy);
            else
                collide_ref(reinterpret_cast<double *>(q), m, om, false, 0.0, 0.0);
        } else if constexpr (W) {
            const MomentsW m{rho - rho0, ux, uy};
            if (force)
                collide_wide(reinterpret_cast<float *>(q), m, rho0, om, true, L.Fx, L.Fy);
            else
                collide_wide(reinterpret_cast<float *>(q), m, rho0, om, false, 0.0, 0.0);
        } else {
            MomentsF m;
            m.dr = (float)(rho - rho0);
            m.ux = (float)ux;
            m.uy = (float)uy;
            if (force)
                collide_f32(reinterpret_cast<float *>(q), m, (float)rho0, (float)om, true,
                            (float)L.Fx, (float)L.Fy);
            else
                collide_f32(reinterpret_cast<float *>(q), m, (float)rho0, (float)om, false, 0.0f,
                            0.0f);
        }
    }
#pragma unroll
    for (int k = 0; k < Q; k++) A[k * P + n] = q[k];
    for (int i = 0; i < L.nprobe; i++)
        if (n == L.probe_n[i]) L.probe_hist[i] = rho;
    if ((L.flags & BF_LINER) && x >= L.gap_x0 && x <= L.gap_x1) {
        const int e = x - L.gap_x0;
        if (y == 0) {
            gptr<double> st = gbl(L.state) + 8 * e;
            for (int i = 0; i < 6; i++) st[i] = 0.0;
            st[6] = rho;
            st[7] = 0.0;
        } else if (y == 1) {
            double *sv = L.saved[0] + 3 * e;
            sv[0] = rho;
            sv[1] = ux;
            sv[2] = uy;
        }
    }
}

// ---- spectra of a finished request's probes (server) ------------------------------
// Block (f, p): P(f) = sum_t p'(t) e^{-2 pi i f t} over samples [t0[p], t1[p]) of
// probe `probe[p]`, p' = cs2 (rho - rho0).  Fixed-order reduction, so the result
// does not depend on anything but the history.
constexpr int kSpecThreads = 256;

struct SpecArgs {
    const double *hist;
    int nprobe;
    int probe[2];
    int t0[2], t1[2];
    double rho0, cs2, f_lo, df;
    double *out;             // nfreq x 2 x (re, im)
};

static __global__ void __launch_bounds__(kSpecThreads) k_spectrum(SpecArgs a)
{
    __shared__ double lre[kSpecThreads], lim[kSpecThreads];
    const int fi = blockIdx.x, p = blockIdx.y;
    const double f = a.f_lo + a.df * fi;
    double re = 0.0, im = 0.0;
    for (int t = a.t0[p] + (int)threadIdx.x; t < a.t1[p]; t += kSpecThreads) {
        const double v = a.cs2 * (a.hist[(long long)t * a.nprobe + a.probe[p]] - a.rho0);
        double sn, cs;
        sincospi(2.0 * f * (double)t, &sn, &cs);
        re += v * cs;
        im -= v * sn;
    }
    lre[threadIdx.x] = re;
    lim[threadIdx.x] = im;
    __syncthreads();
    for (int h = kSpecThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            lre[threadIdx.x] += lre[threadIdx.x + h];
            lim[threadIdx.x] += lim[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.out[4 * fi + 2 * p] = lre[0];
        a.out[4 * fi + 2 * p + 1] = lim[0];
    }
}

}  // namespace lbmx
