// d2q9.h -- D2Q9 lattice tables and the per-node arithmetic of the lbc step,
// shared by every gfx950 kernel of this library.
//
// FP64 routines keep the reference's left-to-right operation order and are
// compiled with -ffp-contract=off, so on IEEE double hardware they round
// exactly like the reference's x86-64 build (no FMA there either):
//   equilibrium      w*rho*(1 + 3cu + 4.5cu^2 - 1.5u^2)     src/core/collision.c:31-32
//   Guo term         w*(3cF + 9cu*cF - 3uF)                  src/core/collision.c:34-35
//   BGK update       f + (-om*(f - feq) + (1 - om/2)*Fk)     src/core/collision.c:37
//   moments          rho = sum f, u = (sum c f + F/2)/rho    compute_macroscopic_fields.c:27-41
// Integer velocities enter as (double)c, exactly as the reference's int*double.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbmx {

constexpr int Q = 9;

struct D2Q9 {
    static constexpr int cx[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static constexpr int cy[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    static constexpr int opp[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
    static constexpr double w[Q] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                                    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

// Runtime-indexed copies for kernels whose direction comes from data (links).
__constant__ static const int c_cx[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
__constant__ static const int c_cy[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
__constant__ static const int c_opp[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
__constant__ static const double c_w[Q] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                                           1.0 / 9.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                           1.0 / 36.0};

// Density floor of compute_macroscopic_fields.c:2,34-36.
constexpr double kRhoFloor = 1e-16;

// Moments of one node in the reference's summation order (k = 0..8).
struct Moments {
    double rho, ux, uy;
};

__device__ __forceinline__ Moments moments_ref(const double f[Q], double Fx, double Fy)
{
    double s = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
    for (int k = 0; k < Q; k++) {
        s += f[k];
        sx += f[k] * (double)D2Q9::cx[k];
        sy += f[k] * (double)D2Q9::cy[k];
    }
    if (s < kRhoFloor) s = kRhoFloor;
    Moments m;
    m.rho = s;
    m.ux = (sx + 0.5 * Fx) / s;
    m.uy = (sy + 0.5 * Fy) / s;
    return m;
}

// Equilibrium of direction k (compile-time k after unrolling, or runtime k via
// the __constant__ tables: identical IEEE results).
__device__ __forceinline__ double feq_ref(double w, double cxk, double cyk, double rho,
                                          double ux, double uy, double usq)
{
    double cu = cxk * ux + cyk * uy;
    return w * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
}

// Post-collision value of one population (collision.c:30-37).  With a zero
// body force the Guo term is +-0 and is skipped: f + (x + -0) == f + x for
// every non-zero x, and the sign of an exact zero population is irrelevant.
__device__ __forceinline__ double collide_one(double f, double w, double cxk, double cyk,
                                              double rho, double ux, double uy, double usq,
                                              double om, bool force, double Fx, double Fy,
                                              double uF)
{
    double cu = cxk * ux + cyk * uy;
    double eq = w * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
    double r = -om * (f - eq);
    if (force) {
        double cF = cxk * Fx + cyk * Fy;
        double Fk = w * (3.0 * cF + 9.0 * cu * cF - 3.0 * uF);
        r = r + (1 - 0.5 * om) * Fk;
    }
    return f + r;
}

__device__ __forceinline__ void collide_ref(double f[Q], const Moments &m, double om,
                                            bool force, double Fx, double Fy)
{
    double usq = m.ux * m.ux + m.uy * m.uy;
    double uF = m.ux * Fx + m.uy * Fy;
#pragma unroll
    for (int k = 0; k < Q; k++)
        f[k] = collide_one(f[k], D2Q9::w[k], (double)D2Q9::cx[k], (double)D2Q9::cy[k], m.rho,
                           m.ux, m.uy, usq, om, force, Fx, Fy, uF);
}

}  // namespace lbmx
