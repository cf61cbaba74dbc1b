// step_kernels.h -- gfx950 kernels of one lbc time step on a resident lattice.
//
// Data layout in HBM (structure of arrays): population k of node (x, y) lives at
// P[k*plane + y*pitch + x], with pitch a multiple of 64 elements and a small
// gap between planes.  Two population buffers S (state at the start of the step,
// the reference's `f`) and D (the reference's `f_new`) swap roles every step.
//
// One step (src/core/main_lbm.c:31-61) is
//   k_bulk      every fluid node: lagged moments, BGK+Guo collision, push of the
//               post-collision populations to the neighbours (streaming.c pull
//               rules), half-way bounce-back into its own slots (bounce_back.c)
//   k_links     solid links: interpolated bounce-back (interpolated_bounce_back.c)
//               and momentum-exchange force partial sums (compute_force_mem.c);
//               k_bulk_links runs both in one launch when S stays untouched
//   k_tail      the boundary table in order (main_lbm.c:49-58), the fixed-order
//               force total and the stale-entry bookkeeping at non-periodic
//               domain edges, in as few launches as their dependences allow
//               (k_bc and k_edges: the same work as stand-alone launches)
// and k_moments produces compute_macroscopic_fields' output when it is asked for.
//
// Semantics reproduced exactly:
//   * collision uses LAGGED moments, i.e. the rho/u/v arrays left by the previous
//     step (collision.c:24-32).  They equal the moments of S for every step after
//     the first, so they are recomputed from S in registers; the first step after
//     an upload reads the uploaded arrays instead.
//   * streaming skips out-of-domain and solid sources (streaming.c:36-40); those
//     D entries keep their old content, which in the reference is the
//     post-collision value of the previous step at the same node.  The push
//     kernel never writes them; k_edges stores that post-collision value into
//     the stale slots of S before S becomes the next step's D.
//   * solid nodes are never written in either buffer.
//   * boundary conditions read lagged moments and post-collision `f`
//     (periodic_x.c) which are recomputed from the untouched S on demand.
// With LBMX_EXACT_STATE the bulk kernel also writes the post-collision values
// in place into S (and the lagged moments into R/U/V), so both buffers hold the
// reference's full state after every step.
#pragma once
#include <math.h>
#include <utility>
#include "d2q9.h"

namespace lbmx {

struct Geo {
    int Nx, Ny, pitch;
    size_t plane;          // elements per population plane (incl. gap)
    int geom;              // 0: all fluid, uniform omega; 1: class byte + omega LUT;
                           // 2: class byte + per-node omega array
    const uint8_t *cls;    // 0 = solid, >= 1 fluid class
    const double *lut;     // omega per class
    const double *om_arr;  // per-node omega (geom 2)
    double om;             // uniform omega (geom 0)
    int per_x, per_y;
    double rho0;           // shift density of FP32 storage
    int wide;              // FP32 storage with FP64 arithmetic (LBMX_F32_WIDE)
};

struct StepFlags {
    int lag_arrays;        // lagged moments come from R/U/V (first step after upload)
    int exact;             // exact-state mode: S holds post-collision values after k_bulk
    int force;             // body force is non-zero
    double Fx, Fy;         // body force of this step
    double Fx_lag, Fy_lag; // body force the lagged moments were computed with
};

struct Fields {
    double *R, *U, *V;     // pitched moment arrays
};

__device__ __forceinline__ bool solid_at(const Geo &g, size_t n)
{
    return g.geom != 0 && g.cls[n] == 0;
}

__device__ __forceinline__ double omega_at(const Geo &g, size_t n)
{
    return g.geom == 0 ? g.om : (g.geom == 1 ? g.lut[g.cls[n]] : g.om_arr[n]);
}

// ---- population storage ----------------------------------------------------
// FP64 stores f itself; FP32-shifted stores g = f - w*rho0 (Lehmann et al.,
// Phys. Rev. E 106, 015308 (2022), "DDF shifting").
__device__ __forceinline__ double pop_get(const double *P, size_t i, int k, double)
{
    return P[i];
}
__device__ __forceinline__ double pop_get(const float *P, size_t i, int k, double rho0)
{
    return (double)P[i] + c_w[k] * rho0;
}
__device__ __forceinline__ void pop_put(double *P, size_t i, int k, double v, double)
{
    P[i] = v;
}
__device__ __forceinline__ void pop_put(float *P, size_t i, int k, double v, double rho0)
{
    P[i] = (float)(v - c_w[k] * rho0);
}

// ---- FP32-shifted collision ------------------------------------------------
// g_eq = feq - w*rho0 = w*(drho + rho*(3cu + 4.5cu^2 - 1.5u^2)), so the O(1)
// part of every population cancels exactly and FP32 carries only deviations.
struct MomentsF {
    float dr, ux, uy;
};

// Explicit fmaf() everywhere a fused multiply-add is wanted (and no contraction
// pragma), so every kernel that inlines these rounds identically.
//
// Nothing here rounds a quantity to FP32's grid around rho0 (spacing ~6e-8 near
// 1).  Such a rounding repeats identically at a slowly changing node step after
// step, so its error accumulates in the velocity instead of averaging out: with
// m.dr = (rho0 + dr) - rho0 and u = j / (rho0 + dr) the velocity drifted from the
// FP64 run by ~3e-8 of u per step on a Taylor-Green vortex.  So dr stays the raw
// sum, and u = (j / rho0)(1 + q) with q = -dr / (rho0 + dr), whose rounding is
// relative to dr, not to rho0.
__device__ __forceinline__ MomentsF moments_f32(const float g[Q], float rho0, float Fx,
                                                float Fy)
{
    const float s56 = g[5] + g[6], s78 = g[7] + g[8];
    float dr = (((g[0] + g[1]) + (g[2] + g[3])) + g[4]) + (s56 + s78);
    const float jx = __fmaf_rn(0.5f, Fx, (g[1] - g[3]) + (g[5] - g[6]) + (g[8] - g[7]));
    const float jy = __fmaf_rn(0.5f, Fy, ((g[2] - g[4]) + s56) - s78);
    float rho = rho0 + dr;
    if (rho < 1e-16f) {
        rho = 1e-16f;
        dr = rho - rho0;
    }
    const float q = -dr * __builtin_amdgcn_rcpf(rho);
    MomentsF m;
    m.dr = dr;
    m.ux = __fmaf_rn(jx, q, jx);
    m.uy = __fmaf_rn(jy, q, jy);
    if (rho0 != 1.0f) {
        m.ux = m.ux / rho0;
        m.uy = m.uy / rho0;
    }
    return m;
}

template <int K>
__device__ __forceinline__ float collide_f32_one(float gk, const MomentsF &m, float rho,
                                                 float usq15, float om, bool force, float Fx,
                                                 float Fy, float uF3, float gf)
{
    constexpr int cx = D2Q9::cx[K], cy = D2Q9::cy[K];
    // the FP32 weights all sit 7.5e-9 above 4/9, 1/9, 1/36: w*X alone would scale
    // the equilibrium's momentum by that much every step, so the product is formed
    // with the weight's remainder and rounded once
    constexpr float w = (float)D2Q9::w[K];
    constexpr float wlo = (float)(D2Q9::w[K] - (double)w);
    float cu = 0.0f, X;
    if constexpr (K == 0) {
        X = __fmaf_rn(rho, -usq15, m.dr);
    } else {
        cu = cdot<cx, cy>(m.ux, m.uy);
        const float t = __fmaf_rn(cu, __fmaf_rn(4.5f, cu, 3.0f), -usq15);
        X = __fmaf_rn(rho, t, m.dr);
    }
    const float geq = __fmaf_rn(w, X, wlo * X);
    float r = __fmaf_rn(-om, gk - geq, gk);
    if (force) {
        float s;
        if constexpr (K == 0) {
            s = -uF3;
        } else {
            const float cF = cdot<cx, cy>(Fx, Fy);
            s = __fmaf_rn(9.0f * cu, cF, __fmaf_rn(3.0f, cF, -uF3));
        }
        r = __fmaf_rn(gf * w, s, r);
    }
    return r;
}

template <int... Ks>
__device__ __forceinline__ void collide_f32_all(float g[Q], const MomentsF &m, float rho,
                                                float usq15, float om, bool force, float Fx,
                                                float Fy, float uF3, float gf,
                                                std::integer_sequence<int, Ks...>)
{
    ((g[Ks] = collide_f32_one<Ks>(g[Ks], m, rho, usq15, om, force, Fx, Fy, uF3, gf)), ...);
}

__device__ __forceinline__ void collide_f32(float g[Q], const MomentsF &m, float rho0, float om,
                                            bool force, float Fx, float Fy)
{
    const float rho = rho0 + m.dr;
    const float usq15 = 1.5f * __fmaf_rn(m.ux, m.ux, m.uy * m.uy);
    const float uF3 = 3.0f * __fmaf_rn(m.ux, Fx, m.uy * Fy);
    const float gf = __fmaf_rn(-0.5f, om, 1.0f);
    collide_f32_all(g, m, rho, usq15, om, force, Fx, Fy, uF3, gf,
                    std::make_integer_sequence<int, Q>{});
}

// ---- FP64 arithmetic on FP32-stored deviations (LBMX_F32_WIDE) ---------------
// Same storage and the same formulas as above, evaluated in FP64 and rounded to
// FP32 once per stored value.  MI355X's FP64 vector rate equals its non-packed
// FP32 rate, so where the step is bound by memory the wider arithmetic costs no
// time, while it removes the FP32 rounding of the equilibrium, whose error (~1e-8
// of g_eq per step, correlated from step to step) otherwise accumulates into the
// slowly varying non-equilibrium part.  Where several steps share one pass the
// instruction count binds, hence the paired, fused form of collide_wide.
struct MomentsW {
    double dr, ux, uy;
};

__device__ __forceinline__ MomentsW moments_wide(const float g[Q], double rho0, double Fx,
                                                 double Fy)
{
    double d[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) d[k] = (double)g[k];
    const double dr = ((d[0] + d[1]) + (d[2] + d[3])) + ((d[4] + d[5]) + (d[6] + d[7])) + d[8];
    const double jx = (d[1] - d[3]) + (d[5] - d[6]) + (d[8] - d[7]);
    const double jy = (d[2] - d[4]) + (d[5] + d[6]) - (d[7] + d[8]);
    double rho = rho0 + dr;
    if (rho < kRhoFloor) rho = kRhoFloor;
    // 1/rho by the hardware reciprocal and two Newton steps (within an ulp of FP64;
    // every value formed from it is rounded to FP32 afterwards) instead of an IEEE
    // division, which costs twice the FP64 instructions
    double inv = __builtin_amdgcn_rcp(rho);
    inv = fma(inv, fma(-rho, inv, 1.0), inv);
    inv = fma(inv, fma(-rho, inv, 1.0), inv);
    MomentsW m;
    m.dr = rho - rho0;
    m.ux = (jx + 0.5 * Fx) * inv;
    m.uy = (jy + 0.5 * Fy) * inv;
    return m;
}

// Opposite directions k and k' = opp(k) see c.u with opposite signs, so their
// equilibria are A +- B with A = w (dr + rho (4.5 (c.u)^2 - 1.5 u.u)) and
// B = 3 w rho (c.u), and their Guo sources 3 c.F + 9 (c.u)(c.F) - 3 u.F are S +- 3 c.F
// with S = 9 (c.u)(c.F) - 3 u.F: one evaluation serves the pair.  Products and
// sums are fused where they meet (fma), which the FP64 reference form does not
// do; here every value is rounded to FP32 afterwards, far above either rounding.
template <int K>
__device__ __forceinline__ void collide_wide_pair(float g[Q], const MomentsW &m, double rho,
                                                  double usq15, double om, bool force, double Fx,
                                                  double Fy, double uF3, double gf)
{
    constexpr int cx = D2Q9::cx[K], cy = D2Q9::cy[K], K2 = D2Q9::opp[K];
    constexpr double w = D2Q9::w[K];
    const double cu = cdot<cx, cy>(m.ux, m.uy);
    const double A = w * fma(rho, fma(4.5 * cu, cu, -usq15), m.dr);
    const double B = (3.0 * w) * (rho * cu);
    const double ga = (double)g[K], gb = (double)g[K2];
    double ra = fma(-om, ga - (A + B), ga);
    double rb = fma(-om, gb - (A - B), gb);
    if (force) {
        const double cF = cdot<cx, cy>(Fx, Fy);
        const double S = fma(9.0 * cu, cF, -uF3);
        const double F3 = 3.0 * cF;
        const double gw = gf * w;
        ra = fma(gw, S + F3, ra);
        rb = fma(gw, S - F3, rb);
    }
    g[K] = (float)ra;
    g[K2] = (float)rb;
}

__device__ __forceinline__ void collide_wide(float g[Q], const MomentsW &m, double rho0, double om,
                                             bool force, double Fx, double Fy)
{
    const double rho = rho0 + m.dr;
    const double usq15 = 1.5 * fma(m.ux, m.ux, m.uy * m.uy);
    const double uF3 = 3.0 * fma(m.ux, Fx, m.uy * Fy);
    const double gf = fma(-0.5, om, 1.0);
    {
        constexpr double w0 = D2Q9::w[0];
        const double g0 = (double)g[0];
        double r = fma(-om, g0 - w0 * fma(rho, -usq15, m.dr), g0);
        if (force) r = fma(gf * w0, -uF3, r);
        g[0] = (float)r;
    }
    collide_wide_pair<1>(g, m, rho, usq15, om, force, Fx, Fy, uF3, gf);   // (1, 3)
    collide_wide_pair<2>(g, m, rho, usq15, om, force, Fx, Fy, uF3, gf);   // (2, 4)
    collide_wide_pair<5>(g, m, rho, usq15, om, force, Fx, Fy, uF3, gf);   // (5, 7)
    collide_wide_pair<8>(g, m, rho, usq15, om, force, Fx, Fy, uF3, gf);   // (8, 6)
}

// Lagged moments in the reference's FP64 form from either FP32 arithmetic.
__device__ __forceinline__ Moments moments_store(const float q[Q], double rho0, double Fx,
                                                 double Fy, bool wide)
{
    if (wide) {
        const MomentsW m = moments_wide(q, rho0, Fx, Fy);
        return Moments{rho0 + m.dr, m.ux, m.uy};
    }
    const MomentsF m = moments_f32(q, (float)rho0, (float)Fx, (float)Fy);
    return Moments{(double)m.dr + rho0, (double)m.ux, (double)m.uy};
}

// ---- lagged moments and post-collision values on demand --------------------
__device__ __forceinline__ Moments lagged_at(const double *S, const Geo &g, const Fields &F,
                                             const StepFlags &s, size_t n)
{
    if (s.lag_arrays || s.exact) return Moments{F.R[n], F.U[n], F.V[n]};
    if (solid_at(g, n)) return Moments{NAN, NAN, NAN};
    double f[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) f[k] = S[k * g.plane + n];
    return moments_ref(f, s.Fx_lag, s.Fy_lag);
}

__device__ __forceinline__ Moments lagged_at(const float *S, const Geo &g, const Fields &F,
                                             const StepFlags &s, size_t n)
{
    if (s.lag_arrays || s.exact) return Moments{F.R[n], F.U[n], F.V[n]};
    if (solid_at(g, n)) return Moments{NAN, NAN, NAN};
    float q[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) q[k] = S[k * g.plane + n];
    return moments_store(q, g.rho0, s.Fx_lag, s.Fy_lag, g.wide);
}

// Post-collision populations of node n (reference values, FP64), i.e. what the
// reference's `f` holds at n after collision.c.  Solid nodes are not collided.
__device__ __forceinline__ void star_node(const double *S, const Geo &g, const Fields &F,
                                          const StepFlags &s, size_t n, double out[Q])
{
#pragma unroll
    for (int k = 0; k < Q; k++) out[k] = S[k * g.plane + n];
    if (s.exact || solid_at(g, n)) return;
    Moments m = s.lag_arrays ? Moments{F.R[n], F.U[n], F.V[n]}
                             : moments_ref(out, s.Fx_lag, s.Fy_lag);
    collide_ref(out, m, omega_at(g, n), s.force, s.Fx, s.Fy);
}

__device__ __forceinline__ void star_node(const float *S, const Geo &g, const Fields &F,
                                          const StepFlags &s, size_t n, double out[Q])
{
    float q[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) q[k] = S[k * g.plane + n];
    if (!(s.exact || solid_at(g, n))) {
        if (g.wide) {
            const MomentsW m = s.lag_arrays ? MomentsW{F.R[n] - g.rho0, F.U[n], F.V[n]}
                                            : moments_wide(q, g.rho0, s.Fx_lag, s.Fy_lag);
            collide_wide(q, m, g.rho0, omega_at(g, n), s.force, s.Fx, s.Fy);
        } else {
            const float r0 = (float)g.rho0;
            MomentsF m;
            if (s.lag_arrays) {
                m.dr = (float)(F.R[n] - g.rho0);
                m.ux = (float)F.U[n];
                m.uy = (float)F.V[n];
            } else {
                m = moments_f32(q, r0, (float)s.Fx_lag, (float)s.Fy_lag);
            }
            collide_f32(q, m, r0, (float)omega_at(g, n), s.force, (float)s.Fx, (float)s.Fy);
        }
    }
#pragma unroll
    for (int k = 0; k < Q; k++) out[k] = (double)q[k] + D2Q9::w[k] * g.rho0;
}

// ---- k_bulk ------------------------------------------------------------------
// Persistent grid: blocks are dealt to the 8 XCDs round-robin, so block b is
// given chunk (b % 8) * (G / 8) + b / 8 and each XCD sweeps one contiguous band
// of rows; within a block, tiles (TX consecutive nodes of one row) are visited in
// row-major order, so the shifted stores of neighbouring tiles meet in one L2.
constexpr int kBulkThreads = 256;

struct BulkArgs {
    Geo g;
    StepFlags s;
    Fields F;
    int tiles_per_row;
    long long ntiles;
    int nxcd;
    int ibb;               // solid links are written by k_links (interpolated BB)
};

// Block b of a persistent bulk grid of G blocks.
template <typename T, int GEOM>
__device__ __forceinline__ void bulk_tiles(T *__restrict__ S, T *__restrict__ D,
                                           const BulkArgs &a, int b, int G)
{
    const Geo &g = a.g;
    const int per_xcd = G / a.nxcd;
    const long long chunk = (long long)(b % a.nxcd) * per_xcd + b / a.nxcd;
    const long long t0 = a.ntiles * chunk / G, t1 = a.ntiles * (chunk + 1) / G;
    const size_t P = g.plane;

    for (long long t = t0; t < t1; t++) {
        const int y = (int)(t / a.tiles_per_row);
        const int x = (int)(t % a.tiles_per_row) * kBulkThreads + threadIdx.x;
        if (x >= g.Nx) continue;
        const size_t n = (size_t)y * g.pitch + x;

        double om;
        if constexpr (GEOM == 0) {
            om = g.om;
        } else {
            const uint8_t c = g.cls[n];
            if (c == 0) {
                if (a.s.exact && !a.s.lag_arrays) {
                    a.F.R[n] = NAN;
                    a.F.U[n] = NAN;
                    a.F.V[n] = NAN;
                }
                continue;
            }
            om = (GEOM == 1) ? g.lut[c] : g.om_arr[n];
        }

        T p[Q];
#pragma unroll
        for (int k = 0; k < Q; k++) p[k] = S[k * P + n];

        if constexpr (sizeof(T) == 8) {
            double *f = reinterpret_cast<double *>(p);
            Moments m;
            if (a.s.lag_arrays) {
                m = Moments{a.F.R[n], a.F.U[n], a.F.V[n]};
            } else {
                m = moments_ref(f, a.s.Fx_lag, a.s.Fy_lag);
                if (a.s.exact) {
                    a.F.R[n] = m.rho;
                    a.F.U[n] = m.ux;
                    a.F.V[n] = m.uy;
                }
            }
            collide_ref(f, m, om, a.s.force, a.s.Fx, a.s.Fy);
        } else if (g.wide) {
            float *q = reinterpret_cast<float *>(p);
            MomentsW m;
            if (a.s.lag_arrays) {
                m = MomentsW{a.F.R[n] - g.rho0, a.F.U[n], a.F.V[n]};
            } else {
                m = moments_wide(q, g.rho0, a.s.Fx_lag, a.s.Fy_lag);
                if (a.s.exact) {
                    a.F.R[n] = g.rho0 + m.dr;
                    a.F.U[n] = m.ux;
                    a.F.V[n] = m.uy;
                }
            }
            collide_wide(q, m, g.rho0, om, a.s.force, a.s.Fx, a.s.Fy);
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
                if (a.s.exact) {
                    a.F.R[n] = (double)m.dr + g.rho0;
                    a.F.U[n] = m.ux;
                    a.F.V[n] = m.uy;
                }
            }
            collide_f32(q, m, r0, (float)om, a.s.force, (float)a.s.Fx, (float)a.s.Fy);
        }

        if (a.s.exact) {
#pragma unroll
            for (int k = 0; k < Q; k++) S[k * P + n] = p[k];
        }

        // push: D[k][x + c_k] = f*_k, wrapping or dropping at the domain edge
#pragma unroll
        for (int k = 0; k < Q; k++) {
            int xd = x + D2Q9::cx[k], yd = y + D2Q9::cy[k];
            long long off = (long long)D2Q9::cy[k] * g.pitch + D2Q9::cx[k];
            if (D2Q9::cx[k] != 0) {
                if (xd < 0) {
                    if (!g.per_x) continue;
                    off += g.Nx;
                } else if (xd >= g.Nx) {
                    if (!g.per_x) continue;
                    off -= g.Nx;
                }
            }
            if (D2Q9::cy[k] != 0) {
                if (yd < 0) {
                    if (!g.per_y) continue;
                    off += (long long)g.Ny * g.pitch;
                } else if (yd >= g.Ny) {
                    if (!g.per_y) continue;
                    off -= (long long)g.Ny * g.pitch;
                }
            }
            const size_t nd = (size_t)((long long)n + off);
            if constexpr (GEOM != 0) {
                if (k != 0 && g.cls[nd] == 0) {
                    // solid destination: half-way bounce-back into the own slot
                    // opp(k) (bounce_back.c:43-46); with interpolated bounce-back
                    // k_links writes that slot instead.
                    if (!a.ibb) D[D2Q9::opp[k] * P + n] = p[k];
                    continue;
                }
            }
            D[k * P + nd] = p[k];
        }
    }
}

template <typename T, int GEOM>
__global__ void __launch_bounds__(kBulkThreads) k_bulk(T *__restrict__ S, T *__restrict__ D,
                                                      BulkArgs a)
{
    bulk_tiles<T, GEOM>(S, D, a, blockIdx.x, gridDim.x);
}

// ---- k_links -----------------------------------------------------------------
// One thread per fluid->solid link (x, k) in row-major order of x, k ascending.
struct Link {
    long long n;    // fluid node
    long long nb;   // behind node x - c_k, or -1 when it lies outside a
                    // non-periodic domain (plain bounce-back then)
    double q;       // wall distance fraction, clamped (interpolated_bounce_back.c:51-61)
    int k;          // direction pointing into the solid
    int bsolid;     // behind node is solid
    int body;       // the solid node x + c_k belongs to the force body (lbmx_set_force_body)
    int pad_;
};

constexpr int kLinkThreads = 256;
constexpr int kPartial = 4;   // doubles per link block: total (x, y), body (x, y)

struct LinkArgs {
    Geo g;
    StepFlags s;
    Fields F;
    const Link *links;
    int nlinks;
    int ibb;
    int write_bb;       // also store plain bounce-back values (stand-alone bounce_back)
    int read_val;       // force only: take f_new(x, opp k) from D instead of computing it
    double *partials;   // kPartial doubles per block
};

// Fixed butterfly within each wavefront, then the waves summed in order; thread 0
// of the block ends with the block sums.
__device__ __forceinline__ void block_sum(double v[kPartial], double *lds)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
#pragma unroll
        for (int c = 0; c < kPartial; c++) v[c] += __shfl_xor(v[c], o, 64);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < kPartial; c++) lds[kPartial * wave + c] = v[c];
    __syncthreads();
    if (threadIdx.x == 0) {
        double s[kPartial] = {0.0, 0.0, 0.0, 0.0};
        for (int w = 0; w < (int)(blockDim.x / 64); w++)
#pragma unroll
            for (int c = 0; c < kPartial; c++) s[c] += lds[kPartial * w + c];
#pragma unroll
        for (int c = 0; c < kPartial; c++) v[c] = s[c];
    }
}

// f*(x, K) and the value (interpolated) bounce-back gives f_new(x, opp K), for a
// link in direction K; K is a compile-time constant so both post-collision node
// arrays stay in registers.
template <int K, typename T>
__device__ __forceinline__ void link_value(const T *__restrict__ S, const LinkArgs &a,
                                           const Link &L, double &fk, double &val)
{
    constexpr int KO = D2Q9::opp[K];
    double st[Q];
    star_node(S, a.g, a.F, a.s, (size_t)L.n, st);
    fk = st[K];
    val = fk;
    if (a.ibb && !a.read_val && L.nb >= 0) {
        const double q = L.q;
        double sb[Q];
        star_node(S, a.g, a.F, a.s, (size_t)L.nb, sb);
        if (q <= 0.5)
            val = (1.0 - 2.0 * q) * fk + 2.0 * q * sb[K];
        else if (!L.bsolid)
            val = (1.0 / (2.0 * q)) * fk + (2.0 * q - 1.0) / (2.0 * q) * sb[KO];
    }
}

// Links [blk * kLinkThreads, +kLinkThreads) and their force partial sum, which goes
// to partials[blk]; bc_force_total sums the partials in a fixed order.
template <typename T>
__device__ __forceinline__ void links_block(const T *__restrict__ S, T *__restrict__ D,
                                            const LinkArgs &a, int blk)
{
    __shared__ double lds[kPartial * kLinkThreads / 64];
    const int l = blk * kLinkThreads + threadIdx.x;
    double f[kPartial] = {0.0, 0.0, 0.0, 0.0};
    if (l < a.nlinks) {
        const Link L = a.links[l];
        const int k = L.k, ko = c_opp[k];
        double fk, val;
        switch (k) {
        case 1: link_value<1>(S, a, L, fk, val); break;
        case 2: link_value<2>(S, a, L, fk, val); break;
        case 3: link_value<3>(S, a, L, fk, val); break;
        case 4: link_value<4>(S, a, L, fk, val); break;
        case 5: link_value<5>(S, a, L, fk, val); break;
        case 6: link_value<6>(S, a, L, fk, val); break;
        case 7: link_value<7>(S, a, L, fk, val); break;
        default: link_value<8>(S, a, L, fk, val); break;
        }
        if (a.read_val) val = pop_get(D, (size_t)ko * a.g.plane + (size_t)L.n, ko, a.g.rho0);
        if ((a.ibb || a.write_bb) && !a.read_val)
            pop_put(D, (size_t)ko * a.g.plane + (size_t)L.n, ko, val, a.g.rho0);
        // compute_force_mem.c:55-57: f*(x,k) + f_new(x,opp k) after (I)BB
        const double df = fk + val;
        f[0] = df * (double)c_cx[k];
        f[1] = df * (double)c_cy[k];
        if (L.body) {
            f[2] = f[0];
            f[3] = f[1];
        }
    }
    block_sum(f, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < kPartial; c++) a.partials[kPartial * blk + c] = f[c];
}

template <typename T>
__global__ void __launch_bounds__(kLinkThreads) k_links(const T *__restrict__ S,
                                                        T *__restrict__ D, LinkArgs a)
{
    links_block(S, D, a, blockIdx.x);
}

// k_bulk and k_links in one launch.  The link blocks come first, so their
// latency-bound work starts at once and overlaps the bulk: blocks [0, nlink) run
// the links, [nlink, link_grid) are padding that keeps the bulk's blocks on the
// XCDs its tile mapping assumes (link_grid is a multiple of the XCD count), and
// the blocks after link_grid run the bulk tiles.  Valid when S is not written during the step (not
// LBMX_EXACT_STATE): the links then read only S, and a link slot D(opp k, n) is
// written either by the bulk (half-way bounce-back, no interpolation) or by the
// link blocks (interpolated bounce-back), never by both -- streaming skips its
// solid source n + c_k -- so the two sets of blocks are independent.
template <typename T, int GEOM>
__global__ void __launch_bounds__(kBulkThreads) k_bulk_links(T *__restrict__ S,
                                                            T *__restrict__ D, BulkArgs ab,
                                                            LinkArgs al, int nlink,
                                                            int link_grid)
{
    static_assert(kBulkThreads == kLinkThreads, "one block size for both parts");
    const int b = blockIdx.x;
    if (b < nlink)
        links_block(S, D, al, b);
    else if (b >= link_grid)
        bulk_tiles<T, GEOM>(S, D, ab, b - link_grid, (int)gridDim.x - link_grid);
}

// ---- the same push with destination-aligned stores -------------------------------
// k_bulk's stores D[k][x + c_k] are shifted by one element for c_k,x != 0: every
// store instruction covers partial cache lines at both ends, which costs most in
// FP32 (4-byte lanes).  Here a workgroup owns the 256 sources x0 .. x0+255 of one
// row y and stores, for each direction k, the destinations x0 .. x0+255 of row
// y + c_k,y -- whole aligned lines -- taking the value for destination x from
// source x - c_k,x: its own lane, a neighbouring lane through LDS, or one of two
// halo sources (x0 - 1, last + 1) whose collision lanes 0 / last compute once more.
// Every slot (destination, k) is written by exactly the one lane that owns it, and
// exactly when k_bulk writes it (source in the domain and fluid, destination
// fluid), with the same value, so the two kernels are interchangeable; the
// half-way bounce-back into a fluid node's own slot (no interpolation) is still
// written by the node's own lane.  Not for LBMX_EXACT_STATE (S stays untouched).
constexpr int kAlignThreads = 256;

// Post-collision populations of fluid node n, as bulk_tiles computes them; W
// selects the FP32-storage collision at compile time (true: FP64 arithmetic,
// g.wide), so each kernel instance carries one collision.
template <typename T, bool W>
__device__ __forceinline__ void collide_loaded(const BulkArgs &a, size_t n, double om, T p[Q])
{
    const Geo &g = a.g;
    if constexpr (sizeof(T) == 8) {
        double *f = reinterpret_cast<double *>(p);
        const Moments m = a.s.lag_arrays ? Moments{a.F.R[n], a.F.U[n], a.F.V[n]}
                                         : moments_ref(f, a.s.Fx_lag, a.s.Fy_lag);
        collide_ref(f, m, om, a.s.force, a.s.Fx, a.s.Fy);
    } else if constexpr (W) {
        float *q = reinterpret_cast<float *>(p);
        const MomentsW m = a.s.lag_arrays ? MomentsW{a.F.R[n] - g.rho0, a.F.U[n], a.F.V[n]}
                                          : moments_wide(q, g.rho0, a.s.Fx_lag, a.s.Fy_lag);
        collide_wide(q, m, g.rho0, om, a.s.force, a.s.Fx, a.s.Fy);
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
        collide_f32(q, m, r0, (float)om, a.s.force, (float)a.s.Fx, (float)a.s.Fy);
    }
}

template <typename T, bool W>
__device__ __forceinline__ void collide_at(const T *__restrict__ S, const BulkArgs &a, size_t n,
                                           double om, T p[Q])
{
#pragma unroll
    for (int k = 0; k < Q; k++) p[k] = S[k * a.g.plane + n];
    collide_loaded<T, W>(a, n, om, p);
}

template <int GEOM>
__device__ __forceinline__ double omega_of(const Geo &g, size_t n)
{
    if constexpr (GEOM == 0) return g.om;
    else if constexpr (GEOM == 1) return g.lut[g.cls[n]];
    else return g.om_arr[n];
}

// LDS slot of direction k with c_k,x = +1 (k = 1, 5, 8 -> 0, 1, 2) or -1
// (k = 3, 6, 7 -> 3, 4, 5)
__device__ __forceinline__ constexpr int xdir_slot(int k)
{
    return k == 1 ? 0 : k == 5 ? 1 : k == 8 ? 2 : k == 3 ? 3 : k == 6 ? 4 : 5;
}

template <typename T, int GEOM, bool W>
__device__ __forceinline__ void aligned_tile(const T *__restrict__ S, T *__restrict__ D,
                                             const BulkArgs &a, int tile)
{
    static_assert(kAlignThreads == kBulkThreads, "row tiles as BulkArgs::tiles_per_row counts them");
    __shared__ T lx[6][kAlignThreads + 2];
    __shared__ int lok[kAlignThreads + 2];
    const Geo &g = a.g;
    const size_t P = g.plane;
    const int y = tile / a.tiles_per_row;
    const int x0 = (tile % a.tiles_per_row) * kAlignThreads;
    const int t = threadIdx.x;
    const int x = x0 + t;
    const int tl = min(x0 + kAlignThreads, g.Nx) - 1 - x0;   // last valid lane
    const bool valid = x < g.Nx;
    const size_t n = (size_t)y * g.pitch + x;

    // halo sources: x0 - 1 (read by lane 0 for c_k,x = +1) into slot 0, and
    // x0 + tl + 1 (read by lane tl for c_k,x = -1) into slot tl + 2.  Lanes 0 and
    // tl load theirs together with their own node, so the two loads overlap; a
    // one-lane tile (tl == 0) takes both halos in lane 0 afterwards.
    const auto halo_source = [&](int side, size_t &ns) -> bool {
        int xs = side == 0 ? x0 - 1 : x0 + tl + 1;
        if (xs < 0 || xs >= g.Nx) {
            if (!g.per_x) return false;
            xs += xs < 0 ? g.Nx : -g.Nx;
        }
        ns = (size_t)y * g.pitch + xs;
        return GEOM == 0 || g.cls[ns] != 0;
    };
    const auto put = [&](int slot, const T v[Q]) {
        lx[0][slot] = v[1];
        lx[1][slot] = v[5];
        lx[2][slot] = v[8];
        lx[3][slot] = v[3];
        lx[4][slot] = v[6];
        lx[5][slot] = v[7];
    };
    const bool one_lane = (tl == 0);
    const bool has_h = !one_lane && (t == 0 || t == tl);
    const int hslot = t == 0 ? 0 : tl + 2;
    size_t hn = 0;
    bool hok = false;
    T h[Q];
    if (has_h) {
        hok = halo_source(t == 0 ? 0 : 1, hn);
        if (hok)
#pragma unroll
            for (int k = 0; k < Q; k++) h[k] = S[k * P + hn];
    }
    T p[Q];
    bool ok = false;
    if (valid) {
        ok = (GEOM == 0) || g.cls[n] != 0;
        if (ok) collide_at<T, W>(S, a, n, omega_of<GEOM>(g, n), p);
        lok[t + 1] = ok;
        if (ok) put(t + 1, p);
    }
    if (has_h) {
        lok[hslot] = hok;
        if (hok) {
            collide_loaded<T, W>(a, hn, omega_of<GEOM>(g, hn), h);
            put(hslot, h);
        }
    }
    if (one_lane && t == 0) {
        for (int side = 0; side < 2; side++) {
            size_t ns = 0;
            const bool sok = halo_source(side, ns);
            const int slot = side == 0 ? 0 : 2;
            lok[slot] = sok;
            if (sok) {
                T v[Q];
                collide_at<T, W>(S, a, ns, omega_of<GEOM>(g, ns), v);
                put(slot, v);
            }
        }
    }
    __syncthreads();
    if (!valid) return;

    // half-way bounce-back into the own slots (no interpolation), as k_bulk
    if (GEOM != 0 && ok && !a.ibb) {
#pragma unroll
        for (int k = 1; k < Q; k++) {
            int xd = x + D2Q9::cx[k], yd = y + D2Q9::cy[k];
            if (xd < 0 || xd >= g.Nx) {
                if (!g.per_x) continue;
                xd += xd < 0 ? g.Nx : -g.Nx;
            }
            if (yd < 0 || yd >= g.Ny) {
                if (!g.per_y) continue;
                yd += yd < 0 ? g.Ny : -g.Ny;
            }
            if (g.cls[(size_t)yd * g.pitch + xd] == 0) D[D2Q9::opp[k] * P + n] = p[k];
        }
    }

#pragma unroll
    for (int k = 0; k < Q; k++) {
        const int cx = D2Q9::cx[k], cy = D2Q9::cy[k];
        int yd = y + cy;
        if (yd < 0 || yd >= g.Ny) {
            if (!g.per_y) continue;          // the whole row of destinations is outside
            yd += yd < 0 ? g.Ny : -g.Ny;
        }
        const size_t nd = (size_t)yd * g.pitch + x;
        bool w;
        T v;
        if (cx == 0) {
            w = ok;
            v = p[k];
        } else {
            const int slot = cx > 0 ? t : t + 2;   // source x - c_k,x
            w = lok[slot];
            v = lx[xdir_slot(k)][slot];
        }
        if (GEOM != 0 && k != 0 && w && g.cls[nd] == 0) w = false;   // solid destination
        if (w) D[k * P + nd] = v;
    }
}

template <typename T, int GEOM, bool W>
__global__ void __launch_bounds__(kAlignThreads) k_bulk_al(T *__restrict__ S, T *__restrict__ D,
                                                          BulkArgs a)
{
    aligned_tile<T, GEOM, W>(S, D, a, blockIdx.x);
}

// k_bulk_al and the links in one launch, link blocks first (as k_bulk_links).
template <typename T, int GEOM, bool W>
__global__ void __launch_bounds__(kAlignThreads) k_bulk_links_al(T *__restrict__ S,
                                                                T *__restrict__ D, BulkArgs ab,
                                                                LinkArgs al, int nlink)
{
    static_assert(kAlignThreads == kLinkThreads, "one block size for both parts");
    const int b = blockIdx.x;
    if (b < nlink)
        links_block(S, D, al, b);
    else
        aligned_tile<T, GEOM, W>(S, D, ab, b - nlink);
}

// ---- k_bc: the boundary table, in order, in one workgroup ----------------------
struct BcDesc {
    int type, index;
    const double *v1, *v2;
};

constexpr int kBcThreads = 1024;

struct BcArgs {
    Geo g;
    StepFlags s;
    Fields F;
    const BcDesc *bcs;
    int nbc;
    const double *partials;
    int npartials;
    double *force_out;   // [0] = x (the reference's D), [1] = y (L); [2], [3] the body's
    double *record;      // optional rows {t, L, D} (cases/cylinder/cylinder.c:162-185)
                         // followed by the body's {L, D}
    long long row;       // row of this step
    int capacity;
    double dt, cf;       // t = row * dt, L and D scaled by cf
};

// Edge nodes of one boundary are independent of each other: a boundary is swept by
// the threads given in a Span (one workgroup in k_bc, several in k_tail), thread
// `first` taking nodes first, first + stride, ...
struct Span {
    int first, stride;
};

template <typename T>
__device__ void bc_zou_he_column(const T *S, T *D, const BcArgs &a, const BcDesc &d, Span sp)
{
    // velocity_inlet.c:28-39, velocity_outlet.c:28-39, pressure_inlet.c:25-39,
    // pressure_outlet.c:25-39
    const Geo &g = a.g;
    const size_t P = g.plane;
    const int i = d.index;
    const bool inlet = (d.type == 7 || d.type == 3);
    const bool pressure = (d.type == 3 || d.type == 4);
    for (int j = sp.first; j < g.Ny; j += sp.stride) {
        const size_t n = (size_t)j * g.pitch + i;
        if (solid_at(g, n)) continue;
        double f[Q];
#pragma unroll
        for (int k = 0; k < Q; k++) f[k] = pop_get(D, k * P + n, k, g.rho0);
        double r, U, V;
        if (inlet) {
            if (!pressure) {
                U = d.v1[j];
                V = d.v2[j];
                r = (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])) / (1.0 - U);
            } else {
                r = d.v1[j];
                U = 1.0 - (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])) / r;
                V = lagged_at(S, g, a.F, a.s, n + 1).uy;
            }
            pop_put(D, 1 * P + n, 1, f[3] + (2.0 / 3.0) * r * U, g.rho0);
            pop_put(D, 5 * P + n, 5,
                    f[7] + (1.0 / 6.0) * r * U - 0.5 * (f[2] - f[4]) + (1.0 / 2.0) * r * V,
                    g.rho0);
            pop_put(D, 8 * P + n, 8,
                    f[6] + (1.0 / 6.0) * r * U + 0.5 * (f[2] - f[4]) - (1.0 / 2.0) * r * V,
                    g.rho0);
        } else {
            if (!pressure) {
                U = d.v1[j];
                V = d.v2[j];
                r = (f[0] + f[2] + f[4] + 2.0 * (f[1] + f[5] + f[8])) / (1.0 + U);
            } else {
                r = d.v1[j];
                U = (f[0] + f[2] + f[4] + 2.0 * (f[1] + f[5] + f[8])) / r - 1.0;
                V = lagged_at(S, g, a.F, a.s, n - 1).uy;
            }
            pop_put(D, 3 * P + n, 3, f[1] - (2.0 / 3.0) * r * U, g.rho0);
            pop_put(D, 6 * P + n, 6,
                    f[8] - 0.5 * (f[2] - f[4]) - (1.0 / 6.0) * r * U + (1.0 / 2.0) * r * V,
                    g.rho0);
            pop_put(D, 7 * P + n, 7,
                    f[5] + 0.5 * (f[2] - f[4]) - (1.0 / 6.0) * r * U - (1.0 / 2.0) * r * V,
                    g.rho0);
        }
    }
}

// ---- per-node boundary closures shared with the batched kernels ---------------
// velocity_top.c:28-39 on one node's post-streaming populations: slots 4, 7, 8.
__device__ __forceinline__ void zou_he_top_node(double f[Q], double U, double V)
{
    const double r = (f[0] + f[1] + f[3] + 2 * (f[2] + f[5] + f[6])) / (1 + V);
    const double f4 = f[2] - 2.0 / 3.0 * r * V;
    const double f7 = f[5] + (f[1] - f[3]) / 2.0 - 0.5 * r * U - 1.0 / 6.0 * r * V;
    const double f8 = f[6] - (f[1] - f[3]) / 2.0 + 0.5 * r * U - 1.0 / 6.0 * r * V;
    f[4] = f4;
    f[7] = f7;
    f[8] = f8;
}

// *_regularized.c:18-71 for one boundary node: fi are the post-streaming
// populations of the neighbouring interior node, (ri, ui, vi) its lagged moments,
// (rb, Ub, Vb) the boundary state; out receives feq(rb, Ub, Vb) + 1.5 w Q:Pi.
__device__ __forceinline__ void regularized_node(const double fi[Q], double ri, double ui,
                                                 double vi, double rb, double Ub, double Vb,
                                                 double out[Q])
{
    const double usq_i = ui * ui + vi * vi;
    const double usq_b = Ub * Ub + Vb * Vb;
    double Pxx = 0.0, Pxy = 0.0, Pyy = 0.0;
#pragma unroll
    for (int k = 0; k < Q; k++) {
        const int cx = D2Q9::cx[k], cy = D2Q9::cy[k];
        double neq = fi[k] - feq_ref(D2Q9::w[k], (double)cx, (double)cy, ri, ui, vi, usq_i);
        Pxx += (double)(cx * cx) * neq;
        Pxy += (double)(cx * cy) * neq;
        Pyy += (double)(cy * cy) * neq;
    }
#pragma unroll
    for (int k = 0; k < Q; k++) {
        const int cx = D2Q9::cx[k], cy = D2Q9::cy[k];
        double eq = feq_ref(D2Q9::w[k], (double)cx, (double)cy, rb, Ub, Vb, usq_b);
        double Qxx = (double)(cx * cx) - 1.0 / 3.0;
        double Qxy = (double)(cx * cy);
        double Qyy = (double)(cy * cy) - 1.0 / 3.0;
        double neq = (3.0 / 2.0) * D2Q9::w[k] * (Qxx * Pxx + 2.0 * Qxy * Pxy + Qyy * Pyy);
        out[k] = eq + neq;
    }
}

template <typename T>
__device__ void bc_zou_he_row(T *D, const BcArgs &a, const BcDesc &d, Span sp)
{
    // velocity_top.c:28-39 and velocity_bottom.c:28-39 (the latter stores its
    // third closure into slot 7, leaving slot 6 as streamed)
    const Geo &g = a.g;
    const size_t P = g.plane;
    const int j = d.index;
    const bool top = (d.type == 9);
    for (int i = sp.first; i < g.Nx; i += sp.stride) {
        const size_t n = (size_t)j * g.pitch + i;
        if (solid_at(g, n)) continue;
        const double U = d.v1[i], V = d.v2[i];
        double f[Q];
#pragma unroll
        for (int k = 0; k < Q; k++) f[k] = pop_get(D, k * P + n, k, g.rho0);
        if (top) {
            zou_he_top_node(f, U, V);
            pop_put(D, 4 * P + n, 4, f[4], g.rho0);
            pop_put(D, 7 * P + n, 7, f[7], g.rho0);
            pop_put(D, 8 * P + n, 8, f[8], g.rho0);
        } else {
            double r = (f[0] + f[1] + f[3] + 2 * (f[4] + f[7] + f[8])) / (1 - V);
            pop_put(D, 2 * P + n, 2, f[4] + 2.0 / 3.0 * r * V, g.rho0);
            pop_put(D, 5 * P + n, 5,
                    f[7] - (f[1] - f[3]) / 2.0 + 0.5 * r * U + 1.0 / 6.0 * r * V, g.rho0);
            pop_put(D, 7 * P + n, 7,
                    f[8] + (f[1] - f[3]) / 2.0 - 0.5 * r * U + 1.0 / 6.0 * r * V, g.rho0);
        }
    }
}

template <typename T>
__device__ void bc_regularized(const T *S, T *D, const BcArgs &a, const BcDesc &d, Span sp)
{
    // src/bc/*_regularized.c:18-71: feq(rho_b, u_b) + 1.5 w Q:Pi with Pi from
    // f_new - feq(lagged) at the neighbouring interior node
    const Geo &g = a.g;
    const size_t P = g.plane;
    const int t = d.type;
    const bool column = (t == 11 || t == 12 || t == 15 || t == 16);
    const bool pressure = (t >= 15 && t <= 18);
    const int step_in = (t == 12 || t == 14 || t == 16 || t == 18) ? 1 : -1;
    const int nedge = column ? g.Ny : g.Nx;
    for (int e = sp.first; e < nedge; e += sp.stride) {
        const int ib = column ? d.index : e, jb = column ? e : d.index;
        const int ii = column ? d.index + step_in : e, ji = column ? e : d.index + step_in;
        const size_t nb = (size_t)jb * g.pitch + ib, ni = (size_t)ji * g.pitch + ii;
        if (solid_at(g, nb)) continue;
        const Moments mi = lagged_at(S, g, a.F, a.s, ni);
        const double ri = mi.rho, ui = mi.ux, vi = mi.uy;
        double rb, Ub, Vb;
        if (pressure) {
            rb = d.v1[e];
            Ub = ui;
            Vb = vi;
        } else {
            rb = ri;
            Ub = d.v1[e];
            Vb = d.v2[e];
        }
        double fi[Q], out[Q];
#pragma unroll
        for (int k = 0; k < Q; k++) fi[k] = pop_get(D, k * P + ni, k, g.rho0);
        regularized_node(fi, ri, ui, vi, rb, Ub, Vb, out);
#pragma unroll
        for (int k = 0; k < Q; k++) pop_put(D, k * P + nb, k, out[k], g.rho0);
    }
}

template <typename T>
__device__ void bc_convective(T *D, const BcArgs &a, const BcDesc &d, Span sp)
{
    // convective_{inlet,outlet,top,bottom}.c:18-28: copy all nine f_new
    const Geo &g = a.g;
    const size_t P = g.plane;
    const int t = d.type;
    const bool column = (t == 1 || t == 2);
    const int from = d.index + ((t == 1 || t == 20) ? 1 : -1);
    const int nedge = column ? g.Ny : g.Nx;
    for (int e = sp.first; e < nedge; e += sp.stride) {
        const size_t nb = column ? (size_t)e * g.pitch + d.index : (size_t)d.index * g.pitch + e;
        const size_t ns = column ? (size_t)e * g.pitch + from : (size_t)from * g.pitch + e;
        if (solid_at(g, nb)) continue;
#pragma unroll
        for (int k = 0; k < Q; k++) D[k * P + nb] = D[k * P + ns];
    }
}

template <typename T>
__device__ void bc_periodic(const T *S, T *D, const BcArgs &a, bool along_x, Span sp)
{
    // periodic_x.c:19-50 / periodic_y.c:19-50: re-pull wrapped populations of the
    // first/last column (row) from post-collision f; the other axis is not wrapped
    const Geo &g = a.g;
    const size_t P = g.plane;
    const int nedge = along_x ? g.Ny : g.Nx;
    const int L = along_x ? g.Nx : g.Ny;
    for (int e = sp.first; e < nedge; e += sp.stride) {
#pragma unroll
        for (int k = 1; k < Q; k++) {
            const int c = along_x ? D2Q9::cx[k] : D2Q9::cy[k];
            if (c == 0) continue;
            const int dst = c > 0 ? 0 : L - 1;
            const int src = c > 0 ? L - c : -c - 1;
            int i, j, is, js;
            if (along_x) {
                i = dst; j = e; is = src; js = e - D2Q9::cy[k];
                if (js < 0 || js >= g.Ny) continue;
            } else {
                i = e; j = dst; is = e - D2Q9::cx[k]; js = src;
                if (is < 0 || is >= g.Nx) continue;
            }
            const size_t nd = (size_t)j * g.pitch + i, ns = (size_t)js * g.pitch + is;
            if (solid_at(g, nd) || solid_at(g, ns)) continue;
            double st[Q];
            star_node(S, g, a.F, a.s, ns, st);
            pop_put(D, k * P + nd, k, st[k], g.rho0);
        }
    }
}

template <typename T>
__device__ __forceinline__ void bc_apply(const T *S, T *D, const BcArgs &a, const BcDesc &d,
                                         Span sp)
{
    switch (d.type) {
    case 3: case 4: case 7: case 8: bc_zou_he_column(S, D, a, d, sp); break;
    case 9: case 10: bc_zou_he_row(D, a, d, sp); break;
    case 1: case 2: case 19: case 20: bc_convective(D, a, d, sp); break;
    case 5: bc_periodic(S, D, a, true, sp); break;
    case 6: bc_periodic(S, D, a, false, sp); break;
    default: bc_regularized(S, D, a, d, sp); break;
    }
}

// main_lbm's force outputs (fixed-order sum of the link partials; force_out[0..1]
// all links as the reference sums them, [2..3] the links into the force body) and
// the record row {t, L, D, L_body, D_body}
constexpr int kRecordCols = 5;

// Called by every thread of one workgroup of NT threads (a power of two): thread t
// sums the partials t, t + NT, ... in that order, then a fixed pairwise tree in LDS
// adds the NT sums, so the result does not depend on timing while all threads'
// loads are in flight together.
template <int NT>
__device__ __forceinline__ void bc_force_total(const BcArgs &a)
{
    static_assert((NT & (NT - 1)) == 0, "power of two");
    __shared__ double red[kPartial][NT];
    const int t = threadIdx.x;
    double s[kPartial] = {0.0, 0.0, 0.0, 0.0};
    for (int p = t; p < a.npartials; p += NT)
#pragma unroll
        for (int c = 0; c < kPartial; c++) s[c] += a.partials[kPartial * p + c];
#pragma unroll
    for (int c = 0; c < kPartial; c++) red[c][t] = s[c];
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int c = 0; c < kPartial; c++) red[c][t] += red[c][t + w];
        __syncthreads();
    }
    if (t != 0) return;
#pragma unroll
    for (int c = 0; c < kPartial; c++) s[c] = red[c][0];
    if (a.npartials > 0)
#pragma unroll
        for (int c = 0; c < kPartial; c++) a.force_out[c] = s[c];
    if (a.record && a.row < a.capacity) {
        double *r = a.record + (size_t)kRecordCols * a.row;
        r[0] = (double)a.row * a.dt;
        r[1] = s[1] * a.cf;
        r[2] = s[0] * a.cf;
        r[3] = s[3] * a.cf;
        r[4] = s[2] * a.cf;
    }
}

// The whole table in one workgroup (stand-alone stages).
template <typename T>
__global__ void __launch_bounds__(kBcThreads) k_bc(const T *__restrict__ S, T *__restrict__ D,
                                                   BcArgs a)
{
    for (int b = 0; b < a.nbc; b++) {
        bc_apply(S, D, a, a.bcs[b], Span{(int)threadIdx.x, (int)blockDim.x});
        __syncthreads();
    }
    bc_force_total<kBcThreads>(a);
}

// ---- k_edges: stale slots of non-periodic domain edges ----------------------------
// The reference leaves D entries whose streaming source is outside the domain
// untouched (streaming.c:36-37); they hold the previous step's post-collision
// value at the same node.  S becomes the next step's D, so its stale slots get
// that value here, after every reader of S in this step has run.
struct EdgeArgs {
    Geo g;
    StepFlags s;
    Fields F;
    int nrow, ncol;   // edge nodes enumerated from rows, then from columns
};

template <typename T>
__device__ __forceinline__ void edge_node(T *__restrict__ S, const EdgeArgs &a, int e)
{
    const Geo &g = a.g;
    if (e >= a.nrow + a.ncol) return;
    int x, y;
    if (e < a.nrow) {   // rows 0 and Ny-1 when y is not periodic
        y = (e < g.Nx) ? 0 : g.Ny - 1;
        x = e % g.Nx;
    } else {            // columns 0 and Nx-1 when x is not periodic (corners excluded
                        // when the rows are already listed)
        const int c = e - a.nrow;
        const int j0 = g.per_y ? 0 : 1;
        const int len = g.per_y ? g.Ny : g.Ny - 2;
        x = (c < len) ? 0 : g.Nx - 1;
        y = j0 + c % len;
    }
    const size_t n = (size_t)y * g.pitch + x;
    if (solid_at(g, n)) return;
    // the lagged moments this step used at n are kept in R/U/V so that the full
    // post-collision field can still be rebuilt later (k_materialize)
    const Moments m = lagged_at(S, g, a.F, a.s, n);
    double st[Q];
    star_node(S, g, a.F, a.s, n, st);
#pragma unroll
    for (int k = 1; k < Q; k++) {
        const int xs = x - D2Q9::cx[k], ys = y - D2Q9::cy[k];
        const bool out = (!g.per_x && (xs < 0 || xs >= g.Nx)) ||
                         (!g.per_y && (ys < 0 || ys >= g.Ny));
        if (out) pop_put(S, k * g.plane + n, k, st[k], g.rho0);
    }
    a.F.R[n] = m.rho;
    a.F.U[n] = m.ux;
    a.F.V[n] = m.uy;
}

template <typename T>
__global__ void __launch_bounds__(256) k_edges(T *__restrict__ S, EdgeArgs a)
{
    edge_node(S, a, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- k_tail: the work after streaming, in as few launches as its order allows ------
// A launch holds jobs -- boundary table entries, or the stale-edge bookkeeping --
// that the host has found independent of each other: no job writes a node another
// job of the launch reads or writes (corners of two boundaries), and the edge job,
// which rewrites S, is only given a launch whose boundaries read S elsewhere.
// Launches run in stream order, so an entry still sees every write of the entries
// before it (main_lbm.c:49-58).  Blocks [block0[j], block0[j+1]) sweep job j; block 0
// of the first launch also forms the force outputs (the link partials are complete
// when it starts).
constexpr int kTailThreads = 128;
constexpr int kTailJobs = 8;

struct TailArgs {
    BcArgs bc;
    int nrow, ncol;                 // the edge job's node counts (EdgeArgs)
    int njobs;
    int with_force;
    int entry[kTailJobs];           // boundary table entry, or -1 for the edge job
    BcDesc desc[kTailJobs];         // that entry's descriptor (no device-table load)
    int block0[kTailJobs + 1];
};

template <typename T>
__global__ void __launch_bounds__(kTailThreads) k_tail(T *__restrict__ S, T *__restrict__ D,
                                                      TailArgs a)
{
    const int blk = blockIdx.x;
    if (a.njobs > 0) {
        int j = 0;
        while (j + 1 < a.njobs && blk >= a.block0[j + 1]) j++;
        const Span sp{(blk - a.block0[j]) * kTailThreads + (int)threadIdx.x,
                      (a.block0[j + 1] - a.block0[j]) * kTailThreads};
        if (a.entry[j] >= 0) {
            bc_apply((const T *)S, D, a.bc, a.desc[j], sp);
        } else {
            const EdgeArgs ea{a.bc.g, a.bc.s, a.bc.F, a.nrow, a.ncol};
            for (int e = sp.first; e < a.nrow + a.ncol; e += sp.stride) edge_node(S, ea, e);
        }
    }
    if (a.with_force && blk == 0) bc_force_total<kTailThreads>(a.bc);   // whole block 0
}

// ---- k_materialize: the reference's post-collision buffer, on demand ----------------
// After a fast-mode step the buffer that was S holds the pre-collision state
// except in the stale slots of edge nodes (already post-collision, k_edges).  This
// kernel finishes the collision in place so the buffer equals the reference's
// `f` after main_lbm: lagged moments from R/U/V where they were kept (edge nodes,
// or every node when the step read the uploaded arrays), else from the buffer.
template <typename T>
__global__ void __launch_bounds__(256) k_materialize(T *__restrict__ S, Geo g, StepFlags s,
                                                     Fields F)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= g.Nx) return;
    const size_t n = (size_t)y * g.pitch + x;
    if (solid_at(g, n)) return;
    bool stale[Q];
    bool edge = false;
#pragma unroll
    for (int k = 0; k < Q; k++) {
        const int xs = x - D2Q9::cx[k], ys = y - D2Q9::cy[k];
        stale[k] = (!g.per_x && (xs < 0 || xs >= g.Nx)) || (!g.per_y && (ys < 0 || ys >= g.Ny));
        edge = edge || stale[k];
    }
    StepFlags t = s;
    t.lag_arrays = s.lag_arrays || edge;   // R/U/V hold the lagged moments here
    t.exact = 0;
    double st[Q];
    star_node(S, g, F, t, n, st);
#pragma unroll
    for (int k = 0; k < Q; k++)
        if (!stale[k]) pop_put(S, k * g.plane + n, k, st[k], g.rho0);
}

// ---- k_moments: compute_macroscopic_fields.c:19-48 --------------------------------
struct MomArgs {
    Geo g;
    Fields F;
    double Fx, Fy;
};

template <typename T>
__global__ void __launch_bounds__(256) k_moments(const T *__restrict__ Dn, MomArgs a)
{
    const Geo &g = a.g;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= g.Nx) return;
    const size_t n = (size_t)y * g.pitch + x;
    double f[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) f[k] = pop_get(Dn, k * g.plane + n, k, g.rho0);
    if (solid_at(g, n)) {
        a.F.R[n] = NAN;
        a.F.U[n] = NAN;
        a.F.V[n] = NAN;
        return;
    }
    const Moments m = moments_ref(f, a.Fx, a.Fy);
    a.F.R[n] = m.rho;
    a.F.U[n] = m.ux;
    a.F.V[n] = m.uy;
}

// ---- device-side Taylor-Green initial state ---------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_init_tgv(T *__restrict__ A, T *__restrict__ B,
                                                  Fields F, Geo g, double u0, double kx,
                                                  double ky)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= g.Nx) return;
    const size_t n = (size_t)y * g.pitch + x;
    const double r = kx / ky;
    const double ux = u0 * sin(kx * x) * cos(ky * y);
    const double uy = -u0 * r * cos(kx * x) * sin(ky * y);
    const double rho = 1.0 - 0.75 * u0 * u0 * (cos(2.0 * kx * x) + r * r * cos(2.0 * ky * y));
    const double usq = ux * ux + uy * uy;
#pragma unroll
    for (int k = 0; k < Q; k++) {
        const double fe = feq_ref(D2Q9::w[k], (double)D2Q9::cx[k], (double)D2Q9::cy[k], rho,
                                  ux, uy, usq);
        pop_put(A, k * g.plane + n, k, fe, g.rho0);
        pop_put(B, k * g.plane + n, k, fe, g.rho0);
    }
    F.R[n] = rho;
    F.U[n] = ux;
    F.V[n] = uy;
}

// ---- layout conversion between the reference's AoS and the planes ----------------
template <typename T>
__global__ void __launch_bounds__(256) k_aos_to_planes(T *__restrict__ Pp,
                                                       const double *__restrict__ aos, Geo g,
                                                       int y0, int rows)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)rows * g.Nx) return;
    const int y = y0 + (int)(t / g.Nx), x = (int)(t % g.Nx);
    const size_t n = (size_t)y * g.pitch + x;
#pragma unroll
    for (int k = 0; k < Q; k++) pop_put(Pp, k * g.plane + n, k, aos[t * Q + k], g.rho0);
}

template <typename T>
__global__ void __launch_bounds__(256) k_planes_to_aos(double *__restrict__ aos,
                                                       const T *__restrict__ Pp, Geo g,
                                                       int y0, int rows)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)rows * g.Nx) return;
    const int y = y0 + (int)(t / g.Nx), x = (int)(t % g.Nx);
    const size_t n = (size_t)y * g.pitch + x;
#pragma unroll
    for (int k = 0; k < Q; k++) aos[t * Q + k] = pop_get(Pp, k * g.plane + n, k, g.rho0);
}

}  // namespace lbmx
