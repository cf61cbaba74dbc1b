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
#include <utility>

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

// c.a for a lattice vector c with components in {-1, 0, 1}: a sum, a difference
// or a signed copy.  The reference multiplies by the components as doubles; a
// product with +-1 is exact and one with 0 only adds a signed zero, so for finite
// values these agree with it bit for bit up to the sign of an exact zero.
template <int CX, int CY, typename T>
__device__ __forceinline__ T cdot(T ax, T ay)
{
    if constexpr (CX == 0) return CY > 0 ? ay : -ay;
    else if constexpr (CY == 0) return CX > 0 ? ax : -ax;
    else if constexpr (CX > 0) return CY > 0 ? ax + ay : ax - ay;
    else return CY > 0 ? ay - ax : -ax - ay;
}

// compute_macroscopic_fields.c:27-41 (and collision.c's lagged moments): the
// sums run over k = 0..8 in order; the terms with a zero velocity component are
// left out of sx and sy (they only add signed zeros).
__device__ __forceinline__ Moments moments_ref(const double f[Q], double Fx, double Fy)
{
    static_assert(D2Q9::cx[1] == 1 && D2Q9::cx[3] == -1 && D2Q9::cx[5] == 1 &&
                      D2Q9::cx[6] == -1 && D2Q9::cx[7] == -1 && D2Q9::cx[8] == 1 &&
                      D2Q9::cy[2] == 1 && D2Q9::cy[4] == -1 && D2Q9::cy[5] == 1 &&
                      D2Q9::cy[6] == 1 && D2Q9::cy[7] == -1 && D2Q9::cy[8] == -1,
                  "D2Q9 ordering of nit.c:195-202");
    double s = f[0];
#pragma unroll
    for (int k = 1; k < Q; k++) s += f[k];
    const double sx = ((((f[1] - f[3]) + f[5]) - f[6]) - f[7]) + f[8];
    const double sy = ((((f[2] - f[4]) + f[5]) + f[6]) - f[7]) - f[8];
    double r = s;
    if (r < kRhoFloor) r = kRhoFloor;
    Moments m;
    m.rho = r;
    m.ux = (sx + 0.5 * Fx) / r;
    m.uy = (sy + 0.5 * Fy) / r;
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

// Post-collision value of population K (collision.c:30-37), the reference's
// expression with c.u and c.F formed by cdot (exact for the +-1 components; the
// zero components only add signed zeros, so for K = 0 the equilibrium is
// w rho (1 - 1.5 u.u) exactly).  With a zero body force the Guo term is +-0 and
// is skipped: f + (x + -0) == f + x for every non-zero x, and the sign of an
// exact zero population is irrelevant.
template <int K>
__device__ __forceinline__ double collide_one(double f, double rho, double ux, double uy,
                                              double usq, double om, bool force, double Fx,
                                              double Fy, double uF)
{
    constexpr int cx = D2Q9::cx[K], cy = D2Q9::cy[K];
    const double w = D2Q9::w[K];
    double cu = 0.0, eq;
    if constexpr (K == 0) {
        eq = w * rho * (1.0 - 1.5 * usq);
    } else {
        cu = cdot<cx, cy>(ux, uy);
        eq = w * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
    }
    double r = -om * (f - eq);
    if (force) {
        double Fk;
        if constexpr (K == 0) {
            Fk = w * -(3.0 * uF);
        } else {
            const double cF = cdot<cx, cy>(Fx, Fy);
            Fk = w * (3.0 * cF + 9.0 * cu * cF - 3.0 * uF);
        }
        r = r + (1 - 0.5 * om) * Fk;
    }
    return f + r;
}

template <int... Ks>
__device__ __forceinline__ void collide_all(double f[Q], const Moments &m, double usq, double om,
                                            bool force, double Fx, double Fy, double uF,
                                            std::integer_sequence<int, Ks...>)
{
    ((f[Ks] = collide_one<Ks>(f[Ks], m.rho, m.ux, m.uy, usq, om, force, Fx, Fy, uF)), ...);
}

__device__ __forceinline__ void collide_ref(double f[Q], const Moments &m, double om,
                                            bool force, double Fx, double Fy)
{
    double usq = m.ux * m.ux + m.uy * m.uy;
    double uF = m.ux * Fx + m.uy * Fy;
    collide_all(f, m, usq, om, force, Fx, Fy, uF, std::make_integer_sequence<int, Q>{});
}

// ---- lattices without a body force ---------------------------------------------
// With Fx = Fy = 0 in every step the reference adds 0.5*F = +0 to the momentum
// sums (compute_macroscopic_fields.c:38-39) and a Guo term of +-0 to each update
// (collision.c:34-37).  Both can change only the sign of an exact zero, and no
// later operation of the step sees that sign: a zero velocity component enters the
// equilibrium only through c.u (then 1 + 3c.u, a sum with 1) and squares, so
// moments_nf / collide_nf give moments_ref / collide_ref's populations bit for bit.
__device__ __forceinline__ Moments moments_nf(const double f[Q])
{
    double s = f[0];
#pragma unroll
    for (int k = 1; k < Q; k++) s += f[k];
    const double sx = ((((f[1] - f[3]) + f[5]) - f[6]) - f[7]) + f[8];
    const double sy = ((((f[2] - f[4]) + f[5]) + f[6]) - f[7]) - f[8];
    double r = s;
    if (r < kRhoFloor) r = kRhoFloor;
    return Moments{r, sx / r, sy / r};
}

// Opposite directions k and opp(k) see c.u = +-c (cdot gives the opposite
// direction's sum as the exact negation), and under round-to-nearest
// 3(-c) = -(3c) and (4.5(-c))(-c) = (4.5c)c exactly, so the reference's
// w rho (1 + 3cu + 4.5cu cu - 1.5u.u) of the pair is formed from one 3c and one
// 4.5c*c: the same roundings on the same operands, with fewer operations.
template <int K>
__device__ __forceinline__ void collide_nf_pair(double f[Q], double c, double wr, double t15,
                                                double nom)
{
    constexpr int K2 = D2Q9::opp[K];
    const double a = 3.0 * c;
    const double b = 4.5 * c * c;
    const double ea = wr * (((1.0 + a) + b) - t15);
    const double eb = wr * (((1.0 - a) + b) - t15);
    f[K] = f[K] + nom * (f[K] - ea);
    f[K2] = f[K2] + nom * (f[K2] - eb);
}

__device__ __forceinline__ void collide_nf(double f[Q], const Moments &m, double om)
{
    const double usq = m.ux * m.ux + m.uy * m.uy;
    const double t15 = 1.5 * usq;
    const double nom = -om;
    const double w0r = D2Q9::w[0] * m.rho, w1r = D2Q9::w[1] * m.rho, w5r = D2Q9::w[5] * m.rho;
    f[0] = f[0] + nom * (f[0] - w0r * (1.0 - t15));
    collide_nf_pair<1>(f, m.ux, w1r, t15, nom);                        // (1, 3): c.u = ux
    collide_nf_pair<2>(f, m.uy, w1r, t15, nom);                        // (2, 4): uy
    collide_nf_pair<5>(f, cdot<1, 1>(m.ux, m.uy), w5r, t15, nom);      // (5, 7): ux + uy
    collide_nf_pair<6>(f, cdot<-1, 1>(m.ux, m.uy), w5r, t15, nom);     // (6, 8): uy - ux
}

}  // namespace lbmx
