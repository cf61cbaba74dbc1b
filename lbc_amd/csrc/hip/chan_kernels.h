// chan_kernels.h -- several steps per launch on bounded lattices: solid walls and
// bodies with half-way or interpolated bounce-back, periodic or closed axes, and
// regularized velocity/pressure columns at x = 0 and x = Nx - 1 (the channel past
// a cylinder of cases/cylinder/cylinder.c and its relatives).
//
// The launch reads the reference's pre-collision field f_k (the one-step push
// path's state) and writes f_{k+D}: row streaming as k_stream (pull_kernels.h),
// one wavefront per level, each level keeping its last R rows in an LDS ring, with
// D + 1 levels for D steps:
//   level 1        loads f_k of its row (aligned, no gather) and collides;
//   levels 2..D    gather f*_{s} of the previous level from LDS, apply the step's
//                  bounce-back links and regularized columns, collide;
//   level D + 1    gathers, applies links and columns, and stores f_{k+D}.
// So every link and column is formed from LDS and registers -- no level waits on a
// chain of dependent global loads -- and what a bounded lattice adds per node and
// step is applied inside the level that computes the node:
//   * a slot whose streaming source is solid takes the (interpolated)
//     bounce-back value of interpolated_bounce_back.c:19-103 / bounce_back.c,
//     formed from the previous level's f* at the node and at its behind node
//     x - c_k, both inside the 3 x 3 neighbourhood the gather reads;
//   * the momentum-exchange term f*(x, k) + f_new(x, opp k) of
//     compute_force_mem.c:47-58 of every link is stored per link and step, and a
//     second kernel sums them in the fixed order of the one-step path (link
//     blocks of 256, then the block partials), so forces do not depend on the
//     launch form;
//   * a regularized column node (*_regularized.c:18-71) takes the neighbouring
//     interior node's post-streaming populations from the adjacent lane (DPP) and
//     that node's lagged moments, which the previous level leaves in a small LDS
//     ring;
//   * solid nodes and nodes outside the domain are computed as garbage and never
//     stored; the reference swaps its two buffers every step, so after an even
//     number of steps the solid nodes' contents of the two buffers are swapped.
// Every value is formed from the same operands with the same operations as the
// one-step push path (k_bulk + k_links + k_tail), so the populations are
// bit-identical to it, in FP64 and in both FP32 storage modes.
//
// Lattices qualify when no entry stays stale (every fluid node with a streaming
// source outside the domain lies on a regularized column, which rewrites all nine
// populations), every regularized node's interior neighbour is fluid, and no
// interpolated link with q <= 1/2 has a solid behind node (the reference reads
// that solid node's buffer contents, which alternate between the two buffers);
// the host checks this before choosing the form.
//
// Kernel arguments hold only what every node needs; the tables of the rare paths
// (link q's, column values, per-class omega, uploaded moments) sit behind one
// pointer and are loaded inside those paths.
#pragma once
#include "pull_kernels.h"

namespace lbmx {

// Tables of the rare paths (device memory, one per lattice).
struct ChanAux {
    const double *lq;         // per link: q, or -1 where the value is plain bounce-back
    int nlinks;
    int lo_type, hi_type;     // regularized column at x = 0 / x = Nx - 1 (0: none)
    const double *lo_v1, *lo_v2, *hi_v1, *hi_v2;
    int geom;                 // 1: omega per class (lut), 2: per node (om_arr)
    const double *lut, *om_arr;
    const double *R, *U, *V;  // lagged moments of the first step after an upload
};

struct ChanArgs {
    int Nx, Ny, pitch, per_x, per_y;
    size_t plane;
    double rho0;
    double om;                // omega of every fluid node when uni
    int uni, wide;
    int force;
    double Fx, Fy;            // body force of every step
    double Fx_lag, Fy_lag;    // force the first step's lagged moments were formed with
    int lag;                  // the first step's lagged moments are the uploaded R/U/V
    int nwork;                // work items (workgroups)
    int lo, hi;               // a regularized column at x = 0 / x = Nx - 1
    // pitched per node: x = class (bits 0-7, 0 = solid) | link bits (bit 7 + k set
    // when x + c_k is solid, k = 1..8); y = index of the node's first link
    const uint2 *node;
    const ChanAux *aux;
    double *ldf;              // [D][nlinks]: f*(x,k) + f_new(x, opp k) of each step
    // per workgroup: first output column, rows [y0, y1), and 0 in .w when its region
    // (strip and rows with the levels' rings) holds only fluid nodes of the uniform
    // omega, away from closed edges and columns
    const int4 *work;
};

// FP64 value of a stored population (pop_get) and back (pop_put).
template <typename T>
__device__ __forceinline__ double chan_load(T v, int k, double rho0)
{
    if constexpr (sizeof(T) == 8) return v;
    else return (double)v + c_w[k] * rho0;
}
template <typename T>
__device__ __forceinline__ T chan_store(double v, int k, double rho0)
{
    if constexpr (sizeof(T) == 8) return v;
    else return (float)(v - c_w[k] * rho0);
}

// Moments and collision of one node as the one-step path forms them, with lagged
// moments m given in the reference's FP64 form (FP32 storage: as its moments).
template <typename T, bool W>
__device__ __forceinline__ Moments chan_moments(const T p[Q], double rho0, double Fx_lag,
                                                double Fy_lag)
{
    if constexpr (sizeof(T) == 8) {
        return moments_ref(reinterpret_cast<const double *>(p), Fx_lag, Fy_lag);
    } else if constexpr (W) {
        const MomentsW m = moments_wide(reinterpret_cast<const float *>(p), rho0, Fx_lag, Fy_lag);
        return Moments{rho0 + m.dr, m.ux, m.uy};
    } else {
        const MomentsF m =
            moments_f32(reinterpret_cast<const float *>(p), (float)rho0, (float)Fx_lag,
                        (float)Fy_lag);
        return Moments{(double)m.dr + rho0, (double)m.ux, (double)m.uy};
    }
}

// Collision with the node's lagged moments: for FP32 storage the moments in FP32
// (or wide) form are recomputed from the populations unless `given` (uploaded
// arrays, converted as k_collide_inplace does).
// NF: no body force in the launch or in the lagged moments; the force terms are
// left out at compile time (the moments keep the reference's sums, so the lagged
// moments a boundary rule reads are unchanged; collide_nf equals collide_ref
// without a force bit for bit).
template <typename T, bool W, bool NF = false>
__device__ __forceinline__ Moments chan_collide(T p[Q], double rho0, double om, double Fx_lag,
                                                double Fy_lag, bool force, double Fx, double Fy,
                                                bool given, Moments mg)
{
    if constexpr (NF) {
        Fx_lag = Fy_lag = Fx = Fy = 0.0;
        force = false;
    }
    if constexpr (sizeof(T) == 8) {
        double *f = reinterpret_cast<double *>(p);
        const Moments m = given ? mg : moments_ref(f, Fx_lag, Fy_lag);
        if constexpr (NF)
            collide_nf(f, m, om);
        else
            collide_ref(f, m, om, force, Fx, Fy);
        return m;
    } else if constexpr (W) {
        float *q = reinterpret_cast<float *>(p);
        const MomentsW m = given ? MomentsW{mg.rho - rho0, mg.ux, mg.uy}
                                 : moments_wide(q, rho0, Fx_lag, Fy_lag);
        collide_wide(q, m, rho0, om, force, Fx, Fy);
        return given ? mg : Moments{rho0 + m.dr, m.ux, m.uy};
    } else {
        float *q = reinterpret_cast<float *>(p);
        const float r0 = (float)rho0;
        MomentsF m;
        if (given) {
            m.dr = (float)(mg.rho - rho0);
            m.ux = (float)mg.ux;
            m.uy = (float)mg.uy;
        } else {
            m = moments_f32(q, r0, (float)Fx_lag, (float)Fy_lag);
        }
        collide_f32(q, m, r0, (float)om, force, (float)Fx, (float)Fy);
        return given ? mg : Moments{(double)m.dr + rho0, (double)m.ux, (double)m.uy};
    }
}

__device__ __forceinline__ double chan_omega(const ChanArgs &a, unsigned code, size_t n)
{
    if (a.uni) return a.om;
    const ChanAux &x = *a.aux;
    return x.geom == 2 ? x.om_arr[n] : x.lut[code & 0xffu];
}

// Bounce-back values of the node's link slots (interpolated_bounce_back.c:51-101;
// with q = -1 the plain value f*(x, k) of bounce_back.c).  star(dx, dy, p) is the
// previous step's f* of population p at (x + dx, y + dy).  With ldf set, each
// link's force term goes to ldf[link].
template <typename T, typename Star>
__device__ __forceinline__ void chan_links(T f[Q], unsigned lm, int base, const double *lq,
                                           double rho0, const Star &star, double *ldf)
{
    int rank = 0;
#pragma unroll
    for (int k = 1; k < Q; k++) {
        if (lm & (1u << k)) {
            const int ko = D2Q9::opp[k];
            const double q = lq[base + rank];
            const double fk = chan_load(star(0, 0, k), k, rho0);
            double val = fk;
            if (q >= 0.0) {
                if (q <= 0.5)
                    val = (1.0 - 2.0 * q) * fk +
                          2.0 * q * chan_load(star(-D2Q9::cx[k], -D2Q9::cy[k], k), k, rho0);
                else
                    val = (1.0 / (2.0 * q)) * fk +
                          (2.0 * q - 1.0) / (2.0 * q) *
                              chan_load(star(-D2Q9::cx[k], -D2Q9::cy[k], ko), ko, rho0);
            }
            f[ko] = chan_store<T>(val, ko, rho0);
            if (ldf) ldf[base + rank] = fk + val;
            rank++;
        }
    }
}

// The regularized closure of one column node from its interior neighbour's
// post-streaming populations fi and lagged moments mi (bc_regularized).
template <typename T>
__device__ __forceinline__ void chan_column(T f[Q], const T fi[Q], const Moments &mi, int type,
                                            double v1, double v2, double rho0)
{
    double rb, Ub, Vb;
    if (type == 15 || type == 16) {   // pressure_{outlet,inlet}_regularized
        rb = v1;
        Ub = mi.ux;
        Vb = mi.uy;
    } else {                          // velocity_{outlet,inlet}_regularized
        rb = mi.rho;
        Ub = v1;
        Vb = v2;
    }
    double di[Q], out[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) di[k] = chan_load(fi[k], k, rho0);
    regularized_node(di, mi.rho, mi.ux, mi.uy, rb, Ub, Vb, out);
#pragma unroll
    for (int k = 0; k < Q; k++) f[k] = chan_store<T>(out[k], k, rho0);
}

template <int NL, int LW, int R>
struct ChanShape {
    static constexpr int TX = LW - 2 * (NL - 1);                 // output columns
    static constexpr int width(int l) { return LW - 2 * (l - 1); }
    static constexpr int offset(int l)
    {
        int o = 0;
        for (int m = 1; m < l; m++) o += R * Q * width(m);
        return o;
    }
    static constexpr int lds = offset(NL);                       // rings of levels 1..NL-1
    static constexpr int mring = (NL - 1) * R * 6;               // moments of columns 1, Nx-2
    static __device__ __forceinline__ int slot(int r, int rbase)
    {
        return (int)((unsigned)(r - rbase + R * NL) % (unsigned)R);
    }
};

// Column of a lane, its periodic wrap or clamp into [0, Nx), inside the domain?
struct ChanPos {
    int x, xw;
    bool xin;
};
__device__ __forceinline__ ChanPos chan_pos(const ChanArgs &a, int x)
{
    ChanPos p;
    p.x = x;
    p.xin = a.per_x || (x >= 0 && x < a.Nx);
    int w = x;
    if (a.per_x) {
        while (w < 0) w += a.Nx;
        while (w >= a.Nx) w -= a.Nx;
    } else {
        w = w < 0 ? 0 : (w >= a.Nx ? a.Nx - 1 : w);
    }
    p.xw = w;
    return p;
}
__device__ __forceinline__ int chan_row(const ChanArgs &a, int r, bool &in)
{
    in = a.per_y || (r >= 0 && r < a.Ny);
    if (a.per_y) {
        while (r < 0) r += a.Ny;
        while (r >= a.Ny) r -= a.Ny;
        return r;
    }
    return r < 0 ? 0 : (r >= a.Ny ? a.Ny - 1 : r);
}

// Level L >= 2: gathers f* of level L-1 from its ring, applies links and columns;
// levels below NL collide into their own ring, level NL stores f_{k+D}.
template <typename T, bool W, int NL, int LW, int P, int R, bool NF, int L>
__device__ __forceinline__ void chan_level(T *buf, double *mring, T *__restrict__ S,
                                           T *__restrict__ Dst, const ChanArgs &a, int lane,
                                           int x0, int y0, int y1, int iters)
{
    using Sh = ChanShape<NL, LW, R>;
    constexpr int Win = Sh::width(L - 1), Wout = Sh::width(L);
    const int rbase = y0 - (NL - 1);
    const T *in = buf + Sh::offset(L - 1);
    const ChanPos Pp = chan_pos(a, x0 - (NL - L) + lane);
    const int xl = x0 - (NL - L);   // column of lane 0
    const bool lo_wave = a.lo && xl <= 0 && xl + Wout > 1;
    const bool hi_wave = a.hi && xl <= a.Nx - 2 && xl + Wout > a.Nx - 1;
    const bool own_x = Pp.x >= x0 && Pp.x < x0 + Sh::TX && Pp.x >= 0 && Pp.x < a.Nx;
    // node words of this level's rows, loaded P rows ahead into the registers they
    // are used from (no copy, so no wait before the barriers)
    auto load_node = [&](int r) -> uint2 {
        bool rin;
        const int rw = chan_row(a, r, rin);
        return (rin && Pp.xin) ? a.node[(size_t)rw * a.pitch + Pp.xw] : make_uint2(0u, 0u);
    };
    // the columns' types and values live in ChanAux; waves holding a column load
    // them once
    int lo_type = 0, hi_type = 0;
    const double *lo_v1 = nullptr, *lo_v2 = nullptr, *hi_v1 = nullptr, *hi_v2 = nullptr;
    if (lo_wave) {
        lo_type = a.aux->lo_type;
        lo_v1 = a.aux->lo_v1;
        lo_v2 = a.aux->lo_v2;
    }
    if (hi_wave) {
        hi_type = a.aux->hi_type;
        hi_v1 = a.aux->hi_v1;
        hi_v2 = a.aux->hi_v2;
    }
    uint2 npre[P];
#pragma unroll
    for (int j = 0; j < P; j++) npre[j] = load_node(rbase - 2 * (L - 1) + j);
    for (int t0 = 0; t0 < iters; t0 += P) {
#pragma unroll
        for (int j = 0; j < P; j++) {
            const int t = t0 + j;
            const int r = rbase + t - 2 * (L - 1);
            const bool active = r >= y0 - (NL - L) && r <= y1 - 1 + (NL - L) && lane < Wout;
            T p[Q];
            Moments m{};
            if (active) {
#pragma unroll
                for (int k = 0; k < Q; k++)
                    p[k] = in[(Sh::slot(r - D2Q9::cy[k], rbase) * Q + k) * Win + lane + 1 -
                              D2Q9::cx[k]];
                bool rin;
                const int rw = chan_row(a, r, rin);
                const unsigned code = npre[j].x;
                const bool own = own_x && r >= y0 && r < y1;
                const unsigned lm = (code >> 7) & 0x1feu;   // bit k: link in direction k
                if (__any(lm != 0u)) {
                    if (lm) {
                        auto star = [&](int dx, int dy, int k) -> T {
                            return in[(Sh::slot(r + dy, rbase) * Q + k) * Win + lane + 1 + dx];
                        };
                        const ChanAux &x_ = *a.aux;
                        chan_links(p, lm, (int)npre[j].y, x_.lq, a.rho0, star,
                                   own ? a.ldf + (size_t)(L - 2) * (size_t)x_.nlinks : nullptr);
                    }
                }
                const bool fluid = (code & 0xffu) != 0u;
                auto mget = [&](int side) -> Moments {
                    const double *mm =
                        mring + (((L - 2) * R + Sh::slot(r, rbase)) * 2 + side) * 3;
                    return Moments{mm[0], mm[1], mm[2]};
                };
                if (lo_wave) {
                    T fi[Q];
#pragma unroll
                    for (int k = 0; k < Q; k++) fi[k] = dpp_shift<kDppShl1>(p[k]);   // x + 1
                    if (Pp.x == 0 && fluid)
                        chan_column(p, fi, mget(0), lo_type, lo_v1[rw], lo_v2 ? lo_v2[rw] : 0.0,
                                    a.rho0);
                }
                if (hi_wave) {
                    T fi[Q];
#pragma unroll
                    for (int k = 0; k < Q; k++) fi[k] = dpp_shift<kDppShr1>(p[k]);   // x - 1
                    if (Pp.x == a.Nx - 1 && fluid)
                        chan_column(p, fi, mget(1), hi_type, hi_v1[rw], hi_v2 ? hi_v2[rw] : 0.0,
                                    a.rho0);
                }
                if constexpr (L < NL) {
                    const size_t n = (size_t)rw * a.pitch + Pp.xw;
                    m = chan_collide<T, W, NF>(p, a.rho0, chan_omega(a, code, n), a.Fx, a.Fy,
                                           a.force, a.Fx, a.Fy, false, Moments{});
                } else if (own) {
                    const size_t o = (size_t)r * a.pitch + Pp.x;
                    if (fluid) {
#pragma unroll
                        for (int k = 0; k < Q; k++) store_nt(&Dst[(size_t)k * a.plane + o], p[k]);
                    } else if ((NL - 1) % 2 == 0) {
                        // an even number of steps: the buffers trade their solid contents
#pragma unroll
                        for (int k = 0; k < Q; k++) {
                            const T s = S[(size_t)k * a.plane + o];
                            S[(size_t)k * a.plane + o] = Dst[(size_t)k * a.plane + o];
                            Dst[(size_t)k * a.plane + o] = s;
                        }
                    }
                }
            }
            if constexpr (L < NL) {
                if constexpr (R == 3) __syncthreads();
                if (active) {
                    T *out = buf + Sh::offset(L) + Sh::slot(r, rbase) * Q * Wout + lane;
#pragma unroll
                    for (int k = 0; k < Q; k++) out[k * Wout] = p[k];
                    double *mm = mring + ((L - 1) * R + Sh::slot(r, rbase)) * 6;
                    if (a.lo && Pp.x == 1) {
                        mm[0] = m.rho;
                        mm[1] = m.ux;
                        mm[2] = m.uy;
                    }
                    if (a.hi && Pp.x == a.Nx - 2) {
                        mm[3] = m.rho;
                        mm[4] = m.ux;
                        mm[5] = m.uy;
                    }
                }
            } else {
                if constexpr (R == 3) __syncthreads();
            }
            npre[j] = load_node(r + P);
            __syncthreads();
        }
    }
}

template <typename T, bool W, int NL, int LW, int P, int R, bool NF, int... Ls>
__device__ __forceinline__ void chan_levels(int level, T *buf, double *mring, T *__restrict__ S,
                                            T *__restrict__ Dst, const ChanArgs &a, int lane,
                                            int x0, int y0, int y1, int iters,
                                            std::integer_sequence<int, Ls...>)
{
    ((level == Ls + 2 ? chan_level<T, W, NL, LW, P, R, NF, Ls + 2>(buf, mring, S, Dst, a, lane, x0,
                                                                y0, y1, iters)
                      : void()),
     ...);
}

// Plain workgroups: every node their levels compute is fluid, relaxes with the
// uniform omega and lies away from closed edges and columns, so the levels run the
// periodic lattice's arithmetic (as k_stream) with no per-node words or checks.
template <typename T, bool W, int NL, int LW, int P, int R, bool NF, int L>
__device__ __forceinline__ void chan_plain_level(T *buf, T *__restrict__ Dst, const ChanArgs &a,
                                                 int lane, int x0, int y0, int y1, int iters)
{
    using Sh = ChanShape<NL, LW, R>;
    constexpr int Win = Sh::width(L - 1), Wout = Sh::width(L);
    const int rbase = y0 - (NL - 1);
    const T *in = buf + Sh::offset(L - 1);
    const int x = x0 - (NL - L) + lane;
    const bool own_x = x >= x0 && x < x0 + Sh::TX && x < a.Nx;
    for (int t = 0; t < iters; t++) {
        const int r = rbase + t - 2 * (L - 1);
        const bool active = r >= y0 - (NL - L) && r <= y1 - 1 + (NL - L) && lane < Wout;
        T p[Q];
        if (active) {
#pragma unroll
            for (int k = 0; k < Q; k++)
                p[k] = in[(Sh::slot(r - D2Q9::cy[k], rbase) * Q + k) * Win + lane + 1 -
                          D2Q9::cx[k]];
            if constexpr (L < NL) {
                chan_collide<T, W, NF>(p, a.rho0, a.om, a.Fx, a.Fy, a.force, a.Fx, a.Fy, false,
                                   Moments{});
            } else if (own_x && r >= y0 && r < y1) {
                char *row = reinterpret_cast<char *>(Dst + (size_t)r * a.pitch);
                const unsigned bx = (unsigned)x * (unsigned)sizeof(T);
#pragma unroll
                for (int k = 0; k < Q; k++)
                    store_nt(reinterpret_cast<T *>(row + (size_t)k * a.plane * sizeof(T) + bx),
                             p[k]);
            }
        }
        if constexpr (R == 3) __syncthreads();
        if constexpr (L < NL) {
            if (active) {
                T *out = buf + Sh::offset(L) + Sh::slot(r, rbase) * Q * Wout + lane;
#pragma unroll
                for (int k = 0; k < Q; k++) out[k * Wout] = p[k];
            }
        }
        __syncthreads();
    }
}

template <typename T, bool W, int NL, int LW, int P, int R, bool NF, int... Ls>
__device__ __forceinline__ void chan_plain_levels(int level, T *buf, T *__restrict__ Dst,
                                                  const ChanArgs &a, int lane, int x0, int y0,
                                                  int y1, int iters,
                                                  std::integer_sequence<int, Ls...>)
{
    ((level == Ls + 2
          ? chan_plain_level<T, W, NL, LW, P, R, NF, Ls + 2>(buf, Dst, a, lane, x0, y0, y1, iters)
          : void()),
     ...);
}

template <typename T, bool W, int NL, int LW, int P, int R, bool NF = false>
__device__ __forceinline__ void chan_plain_first(T *buf, const T *__restrict__ S,
                                                 const ChanArgs &a, int lane, int x0, int y0,
                                                 int y1, int iters)
{
    using Sh = ChanShape<NL, LW, R>;
    const int rbase = y0 - (NL - 1);
    const int n1 = (y1 - y0) + 2 * (NL - 1);
    const ChanPos Pp = chan_pos(a, x0 - (NL - 1) + lane);
    const unsigned bx = (unsigned)Pp.xw * (unsigned)sizeof(T);
    // rows lie within NL - 1 of the lattice: the periodic wrap is one addition
    auto row_ptr = [&](int t) {
        int r = rbase + min(t, n1 - 1);
        r += r < 0 ? a.Ny : 0;
        r -= r >= a.Ny ? a.Ny : 0;
        return reinterpret_cast<const char *>(S + (size_t)r * a.pitch) + bx;
    };
    auto load = [&](int t, T f[Q]) {
        const char *b = row_ptr(t);
#pragma unroll
        for (int k = 0; k < Q; k++)
            f[k] = *reinterpret_cast<const T *>(b + (size_t)k * a.plane * sizeof(T));
    };
    T pre[P][Q];
#pragma unroll
    for (int j = 0; j < P; j++) load(j, pre[j]);
    for (int t0 = 0; t0 < iters; t0 += P) {
#pragma unroll
        for (int j = 0; j < P; j++) {
            const int t = t0 + j;
            if (t < n1)
                chan_collide<T, W, NF>(pre[j], a.rho0, a.om, a.Fx_lag, a.Fy_lag, a.force, a.Fx, a.Fy,
                                   false, Moments{});
            if constexpr (R == 3) __syncthreads();
            if (t < n1) {
                T *out = buf + Sh::slot(rbase + t, rbase) * Q * LW + lane;
#pragma unroll
                for (int k = 0; k < Q; k++) out[k * LW] = pre[j][k];
            }
            load(t + P, pre[j]);
            __syncthreads();
        }
    }
}

// ---- forces of the steps a launch covered ---------------------------------------
// Level l (2..NL) of a launch from f_k stores the momentum-exchange terms of step
// k + l - 2 per link in ldf; k_chan_force forms, per step, the link blocks'
// partial sums (block_sum over 256 links in list order, as links_block), and the
// last block to finish sums the partials as the one-step path's first k_tail
// launch does (bc_force_total with kTailThreads threads), writing the force
// outputs and record rows.  Same terms, same association: the same forces bit for
// bit.
struct ChanForceArgs {
    const Link *links;
    int nlinks;
    const double *ldf;     // [nlev][nlinks]
    int nlev;
    double *partials;      // [nlev][nblk][kPartial]
    int nblk;              // link blocks (0: no links; one block writes zero rows)
    unsigned *done;        // blocks finished (the last one resets it)
    double *force_out;
    double *record;        // rows {t, L, D, L_body, D_body}
    long long row0;
    int capacity;
    double dt, cf;
};

constexpr int kChanMaxLev = 4;   // steps one force launch sums at most

// Lattice direction components from bit masks of k (the d2q9 tables' values:
// c_x = +1 for k = 1, 5, 8 and -1 for 3, 6, 7; c_y = +1 for 2, 5, 6 and -1 for
// 4, 7, 8), so that a lane's direction needs no table load and no branch.
__device__ __forceinline__ int chan_cx(int k)
{
    return (int)((0x122u >> k) & 1u) - (int)((0x0c8u >> k) & 1u);
}
__device__ __forceinline__ int chan_cy(int k)
{
    return (int)((0x064u >> k) & 1u) - (int)((0x190u >> k) & 1u);
}

// v of lane i + 32 (lanes 0 .. 31) or of lane i + 16 (lanes 0 .. 15 and 32 .. 47):
// the second result of v_permlane32_swap / v_permlane16_swap with v as both
// operands (the upper half, or the odd rows, of the first operand swapped with the
// lower half, or the even rows, of the second).
__device__ __forceinline__ double chan_from_upper32(double v)
{
    const unsigned long long b = __double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)b, (unsigned)b, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32),
                                                     false, false);
    return __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}
__device__ __forceinline__ double chan_from_upper16(double v)
{
    const unsigned long long b = __double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)b, (unsigned)b, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(b >> 32), (unsigned)(b >> 32),
                                                     false, false);
    return __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}

// Lane 0's value of the 64-lane xor butterfly of block_sum (v += v of lane ^ o for
// o = 32, 16, ..., 1).  At stage o only lanes below o feed lane 0's later stages,
// and each of them adds the value of lane + o, which it takes by a lane swap (32,
// 16) or a row shift (8 .. 1) with no LDS traffic; every addition is the one the
// butterfly makes, so lane 0 ends with the butterfly's sum bit for bit.  Other
// lanes end with partial values.
__device__ __forceinline__ double chan_wave_sum(double v)
{
    v += chan_from_upper32(v);
    v += chan_from_upper16(v);
    v += dpp_shift<0x108>(v);   // row_shl:8: lane + 8
    v += dpp_shift<0x104>(v);   // row_shl:4
    v += dpp_shift<0x102>(v);   // row_shl:2
    v += dpp_shift<0x101>(v);   // row_shl:1
    return v;
}

// Forces inside the launches, for lattices of at most kChanFusedBlocks link
// blocks: a launch's link terms are summed by the next two launches of the call,
// which need no workgroup to wait for another.  Stage A (in launch k + 1) forms
// every link block's block_sum value per step and component pair into part; stage
// B (in launch k + 2) forms each step's total from them as k_chan_force's last
// block does and writes the force outputs and record rows.
constexpr int kChanFusedBlocks = 128;
constexpr int kChanSplit = 2;   // force workgroups per step and component pair

struct ChanFusedArgs {
    // stage A: terms of the launch before (ldf) -> part_out
    int a_on;
    const Link *links;
    int nlinks, nblk;      // nblk <= kChanFusedBlocks
    const double *ldf;     // [nlev][nlinks]
    double *part_out;      // [nlev][2][kChanFusedBlocks][2]
    // stage B: part_in (stage A of the launch before) -> totals
    int b_on;
    int b_nblk;
    const double *part_in;
    double *force_out;
    double *record;        // rows {t, L, D, L_body, D_body}
    long long row0;
    int capacity;
    double dt, cf;
};

// Stage A in group g (NLEV * 2 * kChanSplit of them): step g / (2 kChanSplit),
// the total's (x, y) components (pair 0) or the body's (pair 1); wave
// j = (g % kChanSplit) * 4 + w of the pair's takes the link blocks b = j,
// j + 4 kChanSplit, ...: the four 64-link butterflies, their sums added in order
// to 0 (links_block's block_sum).
template <int NLEV>
__device__ __forceinline__ void chan_force_partials(const ChanFusedArgs &a, int g)
{
    constexpr int NW = kLinkThreads / 64;
    const int lev = g / (2 * kChanSplit), pair = (g / kChanSplit) % 2;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int stride = kChanSplit * NW;
    const double *ldf = a.ldf + (size_t)lev * a.nlinks;
    double *out = a.part_out + (size_t)(lev * 2 + pair) * kChanFusedBlocks * 2;
    for (int b = (g % kChanSplit) * NW + wave; b < a.nblk; b += stride) {
        int k[NW], body[NW];
        double d[NW];
#pragma unroll
        for (int i = 0; i < NW; i++) {
            // lanes past the last link read the last one and add zeros
            const int l = min(b * kLinkThreads + i * 64 + lane, a.nlinks - 1);
            k[i] = a.links[l].k;
            body[i] = a.links[l].body;
            d[i] = ldf[l];
        }
        double sum[2] = {0.0, 0.0};
#pragma unroll
        for (int i = 0; i < NW; i++) {
            const bool have = b * kLinkThreads + i * 64 + lane < a.nlinks;
            const bool take = have && (pair == 0 || body[i] != 0);
            const double f0 = take ? d[i] * (double)chan_cx(k[i]) : 0.0;
            const double f1 = take ? d[i] * (double)chan_cy(k[i]) : 0.0;
            sum[0] += chan_wave_sum(f0);
            sum[1] += chan_wave_sum(f1);
        }
        if (lane == 0) {
            out[2 * b] = sum[0];
            out[2 * b + 1] = sum[1];
        }
    }
}

// Stage B of step lev, pair (wave 0): lane t adds the block values t, t + 64 (each
// to 0, as bc_force_total<128>'s loops do), then shuffles down by 32 .. 1:
// k_chan_force's association, the one-step path's sums.
template <int NLEV>
__device__ __forceinline__ void chan_force_totals(const ChanFusedArgs &a, int lev, int pair)
{
    static_assert(kChanFusedBlocks == 2 * 64, "lane t sums blocks t and t + 64");
    const int lane = threadIdx.x;
    const double *in = a.part_in + (size_t)(lev * 2 + pair) * kChanFusedBlocks * 2;
    double v[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        double sa = 0.0, sb = 0.0;
        if (lane < a.b_nblk) sa += in[2 * lane + c];
        if (lane + 64 < a.b_nblk) sb += in[2 * (lane + 64) + c];
        v[c] = sa + sb;
    }
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1)
#pragma unroll
        for (int c = 0; c < 2; c++) v[c] += __shfl_down(v[c], w, 64);
    if (lane != 0) return;
    if (a.b_nblk > 0 && lev == NLEV - 1) {
        a.force_out[2 * pair] = v[0];
        a.force_out[2 * pair + 1] = v[1];
    }
    const long long row = a.row0 + lev;
    if (a.record && row < a.capacity) {
        double *r = a.record + (size_t)kRecordCols * row;
        if (pair == 0) r[0] = (double)row * a.dt;
        r[1 + 2 * pair] = v[1] * a.cf;
        r[2 + 2 * pair] = v[0] * a.cf;
    }
}

// NL - 1 steps per launch; S holds f_k (only its solid nodes are written), Dst
// receives f_{k+NL-1} at the fluid nodes.  With fj.a_on or fj.b_on the first
// 2 (NL - 1) kChanSplit blocks are force workgroups: group g runs stage B of its
// step and pair when g % kChanSplit == 0, then its share of stage A; the work
// items follow.
template <typename T, bool W, int NL, int LW, int P, int R, bool NF = false>
__global__ void __launch_bounds__(LW * NL) __attribute__((amdgpu_waves_per_eu(4)))
k_chan(T *__restrict__ S, T *__restrict__ Dst, ChanArgs a, ChanFusedArgs fj)
{
    using Sh = ChanShape<NL, LW, R>;
    static_assert(LW == 64 && NL >= 2 && Sh::TX > 0 && (R == 3 || R == 4), "strip shape");
    static_assert(LW * NL == kLinkThreads, "a force workgroup is one link block wide");
    __shared__ T buf[Sh::lds];
    __shared__ double mring[Sh::mring];
    const int nf = (fj.a_on || fj.b_on) ? 2 * (NL - 1) * kChanSplit : 0;
    if ((int)blockIdx.x < nf) {
        const int g = blockIdx.x;
        if (fj.b_on && g % kChanSplit == 0 && threadIdx.x < 64)
            chan_force_totals<NL - 1>(fj, g / (2 * kChanSplit), (g / kChanSplit) % 2);
        if (fj.a_on) chan_force_partials<NL - 1>(fj, g);
        return;
    }
    const int level = __builtin_amdgcn_readfirstlane(threadIdx.x / LW) + 1;
    const int lane = threadIdx.x % LW;
    // XCD-aware order: consecutive workgroups go to the 8 XCDs in turn, so each XCD
    // takes one eighth of the work list (strips of one row band, then the next)
    const int bw = (int)blockIdx.x - nf;
    const int per = (a.nwork + 7) / 8;
    const int item = (bw % 8) * per + bw / 8;
    if (item >= a.nwork) return;   // whole workgroup, before any barrier
    const int4 wk = a.work[item];
    const int x0 = wk.x, y0 = wk.y, y1 = wk.z;
    const int iters = ((y1 - y0) + 3 * (NL - 1) + P - 1) / P * P;
    const int rbase = y0 - (NL - 1);

    if (!a.lag && wk.w == 0) {
        if (level == 1)
            chan_plain_first<T, W, NL, LW, P, R, NF>(buf, S, a, lane, x0, y0, y1, iters);
        else
            chan_plain_levels<T, W, NL, LW, P, R, NF>(level, buf, Dst, a, lane, x0, y0, y1, iters,
                                                  std::make_integer_sequence<int, NL - 1>{});
        return;
    }
    if (level == 1) {
        // load f_k of the row (own node, aligned) and collide
        const int n1 = (y1 - y0) + 2 * (NL - 1);
        const ChanPos Pp = chan_pos(a, x0 - (NL - 1) + lane);
        auto row = [&](int t) {
            bool in;
            return chan_row(a, rbase + min(t, n1 - 1), in);
        };
        auto load_row = [&](int t, T f[Q], uint2 &nd) {
            bool in;
            const int rr = chan_row(a, rbase + min(t, n1 - 1), in);
            const size_t n = (size_t)rr * a.pitch + Pp.xw;
#pragma unroll
            for (int k = 0; k < Q; k++) f[k] = S[(size_t)k * a.plane + n];
            nd = (in && Pp.xin) ? a.node[n] : make_uint2(0u, 0u);
        };
        T pre[P][Q];
        uint2 npre[P];
#pragma unroll
        for (int j = 0; j < P; j++) load_row(j, pre[j], npre[j]);
        for (int t0 = 0; t0 < iters; t0 += P) {
#pragma unroll
            for (int j = 0; j < P; j++) {
                const int t = t0 + j;
                const int r = rbase + t;
                Moments m{};
                if (t < n1) {
                    const size_t n = (size_t)row(t) * a.pitch + Pp.xw;
                    Moments mg{};
                    if (a.lag) mg = Moments{a.aux->R[n], a.aux->U[n], a.aux->V[n]};
                    m = chan_collide<T, W, NF>(pre[j], a.rho0, chan_omega(a, npre[j].x, n), a.Fx_lag,
                                           a.Fy_lag, a.force, a.Fx, a.Fy, a.lag, mg);
                }
                if constexpr (R == 3) __syncthreads();
                if (t < n1) {
                    T *out = buf + Sh::slot(r, rbase) * Q * LW + lane;
#pragma unroll
                    for (int k = 0; k < Q; k++) out[k * LW] = pre[j][k];
                    double *mm = mring + Sh::slot(r, rbase) * 6;
                    if (a.lo && Pp.x == 1) {
                        mm[0] = m.rho;
                        mm[1] = m.ux;
                        mm[2] = m.uy;
                    }
                    if (a.hi && Pp.x == a.Nx - 2) {
                        mm[3] = m.rho;
                        mm[4] = m.ux;
                        mm[5] = m.uy;
                    }
                }
                load_row(t + P, pre[j], npre[j]);
                __syncthreads();
            }
        }
    } else {
        chan_levels<T, W, NL, LW, P, R, NF>(level, buf, mring, S, Dst, a, lane, x0, y0, y1, iters,
                                        std::make_integer_sequence<int, NL - 1>{});
    }
}

// Per step, thread t of block b takes link b * 256 + t; the step's block partial
// is links_block's block_sum (a butterfly within each wavefront, then the
// wavefronts' sums added in order by one thread), all steps at once with one
// barrier.  The last block to finish forms each step's total in one wavefront per
// step with bc_force_total<kTailThreads>'s association: lane t sums the partials
// of threads t and t + 64 (p = t, t + 128, ... in order), the first tree stage adds
// the two, and the remaining stages (w = 32 .. 1: red[t] += red[t + w]) become
// shuffles down by w.  The same terms in the same association: the one-step
// path's forces bit for bit.
__global__ void __launch_bounds__(kLinkThreads) k_chan_force(ChanForceArgs a)
{
    static_assert(kLinkThreads == 256 && kTailThreads == 128, "association of the one-step path");
    constexpr int NW = kLinkThreads / 64;
    __shared__ double lds[kChanMaxLev][NW][kPartial];
    __shared__ int last;
    const int l = blockIdx.x * kLinkThreads + threadIdx.x;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (a.nblk > 0) {
        const bool have = l < a.nlinks;
        int k = 0, body = 0;
        double d[kChanMaxLev];
        if (have) {
            const Link L = a.links[l];
            k = L.k;
            body = L.body;
        }
#pragma unroll
        for (int lev = 0; lev < kChanMaxLev; lev++)
            d[lev] = (have && lev < a.nlev) ? a.ldf[(size_t)lev * a.nlinks + l] : 0.0;
#pragma unroll
        for (int lev = 0; lev < kChanMaxLev; lev++) {
            if (lev >= a.nlev) break;
            double f[kPartial] = {0.0, 0.0, 0.0, 0.0};
            if (have) {
                f[0] = d[lev] * (double)c_cx[k];
                f[1] = d[lev] * (double)c_cy[k];
                if (body) {
                    f[2] = f[0];
                    f[3] = f[1];
                }
            }
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1)
#pragma unroll
                for (int c = 0; c < kPartial; c++) f[c] += __shfl_xor(f[c], o, 64);
            if (lane == 0)
#pragma unroll
                for (int c = 0; c < kPartial; c++) lds[lev][wave][c] = f[c];
        }
        __syncthreads();
        if (threadIdx.x < a.nlev) {
            const int lev = threadIdx.x;
            double s[kPartial] = {0.0, 0.0, 0.0, 0.0};
            for (int w = 0; w < NW; w++)
#pragma unroll
                for (int c = 0; c < kPartial; c++) s[c] += lds[lev][w][c];
#pragma unroll
            for (int c = 0; c < kPartial; c++)
                a.partials[((size_t)lev * a.nblk + blockIdx.x) * kPartial + c] = s[c];
        }
        // the last block to finish forms the totals
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) last = atomicAdd(a.done, 1u) == (unsigned)(gridDim.x - 1);
        __syncthreads();
        if (!last) return;
        __threadfence();
    }
    if (wave >= a.nlev) return;
    const int lev = wave;
    const double *part = a.partials + (size_t)lev * a.nblk * kPartial;
    double v[kPartial];
#pragma unroll
    for (int c = 0; c < kPartial; c++) {
        double sa = 0.0, sb = 0.0;
        for (int p = lane; p < a.nblk; p += kTailThreads) sa += part[kPartial * p + c];
        for (int p = lane + 64; p < a.nblk; p += kTailThreads) sb += part[kPartial * p + c];
        v[c] = sa + sb;
    }
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1)
#pragma unroll
        for (int c = 0; c < kPartial; c++) v[c] += __shfl_down(v[c], w, 64);
    if (lane != 0) return;
    if (a.nblk > 0 && lev == a.nlev - 1)
#pragma unroll
        for (int c = 0; c < kPartial; c++) a.force_out[c] = v[c];
    const long long row = a.row0 + lev;
    if (a.record && row < a.capacity) {
        double *r = a.record + (size_t)kRecordCols * row;
        r[0] = (double)row * a.dt;
        r[1] = v[1] * a.cf;
        r[2] = v[0] * a.cf;
        r[3] = v[3] * a.cf;
        r[4] = v[2] * a.cf;
    }
    if (lev == 0 && a.nblk > 0) *a.done = 0u;
}

}  // namespace lbmx
