/*
 * lbm_oracle.c -- CPU restatement of the lbc D2Q9 BGK step (TEST INFRASTRUCTURE).
 *
 * This file is the checker for the HIP path, not part of the product.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each routine restates one reference routine (file:line given per function),
 * keeping the reference's loop order and the left-to-right floating-point
 * evaluation order of its C expressions.  Built with gcc and no FMA contraction
 * (x86-64 baseline, as the reference's own Makefile builds), results are
 * bit-identical to the reference.  The only order-dependent quantity is the
 * momentum-exchange force, which the reference sums with OpenMP atomics in
 * thread order; here it is a serial row-major sum.
 *
 * Integer lattice vectors enter the arithmetic exactly as in the reference:
 * (double)c * x, so products with c in {-1,0,1} are exact.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "lbm_oracle.h"

#define Q 9
static const int CX[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int CY[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int OPP[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
static const double W[Q] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                            1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

#define NODE(i, j) ((size_t)(j) * (size_t)Nx + (size_t)(i))
#define POP(i, j, k) (NODE(i, j) * Q + (size_t)(k))

/* Second-order equilibrium as written in the reference's collision and BCs:
 * w*rho*(1 + 3cu + 4.5cu^2 - 1.5u^2), evaluated left to right. */
static inline double feq_c(int k, double rho, double ux, double uy, double usq)
{
    double cu = CX[k] * ux + CY[k] * uy;
    return W[k] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
}

/* Wraps (periodic) or rejects (returns 0) a neighbour coordinate, as the
 * neighbour tests in bounce_back.c:26-40 and compute_force_mem.c:35-49 do. */
static inline int wrap_or_reject(int *c, int n, int periodic)
{
    if (periodic) {
        if (*c < 0) *c += n;
        else if (*c >= n) *c -= n;
        return 1;
    }
    return (*c >= 0 && *c < n);
}

/* src/core/collision.c:18-40 -- in-place BGK with Guo forcing, lagged moments. */
void orc_collision(int Nx, int Ny, double *f, const double *rho, const double *u,
                   const double *v, const double *omega, const int *mask,
                   double Fx, double Fy)
{
    #pragma omp parallel for
    for (int i = 0; i < Nx; i++) {
        for (int j = 0; j < Ny; j++) {
            size_t n = NODE(i, j);
            if (mask[n]) continue;
            double ux = u[n], uy = v[n];
            double usq = ux * ux + uy * uy;
            double uF = ux * Fx + uy * Fy;
            double om = omega[n];
            for (int k = 0; k < Q; k++) {
                double cu = CX[k] * ux + CY[k] * uy;
                double eq = W[k] * rho[n] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
                double cF = CX[k] * Fx + CY[k] * Fy;
                double Fk = W[k] * (3.0 * cF + 9.0 * cu * cF - 3.0 * uF);
                double *p = &f[n * Q + k];
                *p += -om * (*p - eq) + (1 - 0.5 * om) * Fk;
            }
        }
    }
}

/* src/core/streaming.c:14-44 -- pull; out-of-domain and solid sources are left
 * untouched in f_new (their old content survives). */
void orc_streaming(int Nx, int Ny, const double *f, double *f_new, const int *mask,
                   int per_x, int per_y)
{
    #pragma omp parallel for
    for (int i = 0; i < Nx; i++) {
        for (int j = 0; j < Ny; j++) {
            if (mask[NODE(i, j)]) continue;
            for (int k = 0; k < Q; k++) {
                int is = i - CX[k], js = j - CY[k];
                if (per_x) { if (is < 0) is += Nx; else if (is >= Nx) is -= Nx; }
                if (per_y) { if (js < 0) js += Ny; else if (js >= Ny) js -= Ny; }
                if (is < 0 || is >= Nx || js < 0 || js >= Ny) continue;
                if (mask[NODE(is, js)]) continue;
                f_new[POP(i, j, k)] = f[POP(is, js, k)];
            }
        }
    }
}

/* src/bc/bounce_back.c:15-49 -- half-way bounce-back written before streaming. */
void orc_bounce_back(int Nx, int Ny, const double *f, double *f_new, const int *mask,
                     int per_x, int per_y)
{
    #pragma omp parallel for
    for (int i = 0; i < Nx; i++) {
        for (int j = 0; j < Ny; j++) {
            if (mask[NODE(i, j)]) continue;
            for (int k = 0; k < Q; k++) {
                int in = i + CX[k], jn = j + CY[k];
                if (!wrap_or_reject(&in, Nx, per_x)) continue;
                if (!wrap_or_reject(&jn, Ny, per_y)) continue;
                if (!mask[NODE(in, jn)]) continue;
                f_new[POP(i, j, OPP[k])] = f[POP(i, j, k)];
            }
        }
    }
}

/* src/bc/interpolated_bounce_back.c:19-103.  The link fraction q comes from the
 * signed distance phi and is clamped to [1e-12, 1-1e-12]; the interpolation
 * uses the behind node x - c_k as the reference does (its q <= 1/2 branch mixes
 * f*(x,k) with f*(x-c_k,k); its q > 1/2 branch uses f*(x-c_k, opp k)). */
void orc_interpolated_bounce_back(int Nx, int Ny, const double *f, double *f_new,
                                  const int *mask, const double *phi, int per_x, int per_y)
{
    const double eps = 1e-12;
    #pragma omp parallel for
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) {
            if (mask[NODE(i, j)]) continue;
            for (int k = 0; k < Q; k++) {
                int in = i + CX[k], jn = j + CY[k];
                if (!wrap_or_reject(&in, Nx, per_x)) continue;
                if (!wrap_or_reject(&jn, Ny, per_y)) continue;
                if (!mask[NODE(in, jn)]) continue;

                double den = phi[NODE(i, j)] - phi[NODE(in, jn)];
                if (fabs(den) < eps) den = (den >= 0.0) ? eps : -eps;
                double q = phi[NODE(i, j)] / den;
                if (q < eps) q = eps;
                if (q > 1.0 - eps) q = 1.0 - eps;

                size_t dst = POP(i, j, OPP[k]);
                double fk = f[POP(i, j, k)];
                int ib = i - CX[k], jb = j - CY[k];
                if (!wrap_or_reject(&ib, Nx, per_x) || !wrap_or_reject(&jb, Ny, per_y)) {
                    f_new[dst] = fk;   /* no behind node: plain bounce-back */
                    continue;
                }
                if (q <= 0.5) {
                    f_new[dst] = (1.0 - 2.0 * q) * fk + 2.0 * q * f[POP(ib, jb, k)];
                } else if (mask[NODE(ib, jb)]) {
                    f_new[dst] = fk;
                } else {
                    f_new[dst] = (1.0 / (2.0 * q)) * fk
                               + (2.0 * q - 1.0) / (2.0 * q) * f[POP(ib, jb, OPP[k])];
                }
            }
        }
    }
}

/* src/core/compute_force_mem.c:19-71 -- momentum exchange over every
 * fluid->solid link (walls included); serial row-major order. */
void orc_force_mem(int Nx, int Ny, const double *f, const double *f_new, const int *mask,
                   int per_x, int per_y, double *Fx_out, double *Fy_out)
{
    double fx = 0.0, fy = 0.0;
    for (int j = 0; j < Ny; ++j) {
        for (int i = 0; i < Nx; ++i) {
            if (mask[NODE(i, j)]) continue;
            for (int k = 1; k < Q; ++k) {
                int in = i + CX[k], jn = j + CY[k];
                if (!wrap_or_reject(&in, Nx, per_x)) continue;
                if (!wrap_or_reject(&jn, Ny, per_y)) continue;
                if (mask[NODE(in, jn)]) {
                    double df = f[POP(i, j, k)] + f_new[POP(i, j, OPP[k])];
                    fx += df * CX[k];
                    fy += df * CY[k];
                }
            }
        }
    }
    *Fx_out = fx;
    *Fy_out = fy;
}

/* src/core/compute_macroscopic_fields.c:19-48 -- density clamp at 1e-16, Guo
 * half-force velocity shift, NaN on solid nodes. */
void orc_moments(int Nx, int Ny, const double *f_new, const int *mask, double *rho,
                 double *u, double *v, double Fx, double Fy)
{
    #pragma omp parallel for
    for (int i = 0; i < Nx; i++) {
        for (int j = 0; j < Ny; j++) {
            size_t n = NODE(i, j);
            double s = 0.0, sx = 0.0, sy = 0.0;
            for (int k = 0; k < Q; k++) {
                double fk = f_new[n * Q + k];
                s += fk;
                sx += fk * CX[k];
                sy += fk * CY[k];
            }
            if (s < 1e-16) s = 1e-16;
            if (!mask[n]) {
                rho[n] = s;
                u[n] = (sx + 0.5 * Fx) / s;
                v[n] = (sy + 0.5 * Fy) / s;
            } else {
                rho[n] = NAN;
                u[n] = NAN;
                v[n] = NAN;
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * Boundary conditions (src/bc/ files).  Each runs over one column (inlet/outlet)
 * or one row (top/bottom) and skips solid boundary nodes.
 * ------------------------------------------------------------------------- */

/* Zou-He velocity/pressure closures on a column.  src/bc/velocity_inlet.c:28-39,
 * velocity_outlet.c:28-39, pressure_inlet.c:25-39, pressure_outlet.c:25-39. */
static void zou_he_column(int Nx, int Ny, int type, int i, const double *val1,
                          const double *val2, double *fn, const double *v, const int *mask)
{
    for (int j = 0; j < Ny; j++) {
        if (mask[NODE(i, j)]) continue;
        double *p = &fn[POP(i, j, 0)];
        double f0 = p[0], f1 = p[1], f2 = p[2], f3 = p[3], f4 = p[4];
        double f5 = p[5], f6 = p[6], f7 = p[7], f8 = p[8];
        double r, U, V;
        if (type == ORC_BC_VELOCITY_INLET || type == ORC_BC_PRESSURE_INLET) {
            if (type == ORC_BC_VELOCITY_INLET) {
                U = val1[j];
                V = val2[j];
                r = (f0 + f2 + f4 + 2.0 * (f3 + f6 + f7)) / (1.0 - U);
            } else {
                r = val1[j];
                U = 1.0 - (f0 + f2 + f4 + 2.0 * (f3 + f6 + f7)) / r;
                V = v[NODE(i + 1, j)];
            }
            p[1] = f3 + (2.0 / 3.0) * r * U;
            p[5] = f7 + (1.0 / 6.0) * r * U - 0.5 * (f2 - f4) + (1.0 / 2.0) * r * V;
            p[8] = f6 + (1.0 / 6.0) * r * U + 0.5 * (f2 - f4) - (1.0 / 2.0) * r * V;
        } else {
            if (type == ORC_BC_VELOCITY_OUTLET) {
                U = val1[j];
                V = val2[j];
                r = (f0 + f2 + f4 + 2.0 * (f1 + f5 + f8)) / (1.0 + U);
            } else {
                r = val1[j];
                U = (f0 + f2 + f4 + 2.0 * (f1 + f5 + f8)) / r - 1.0;
                V = v[NODE(i - 1, j)];
            }
            p[3] = f1 - (2.0 / 3.0) * r * U;
            p[6] = f8 - 0.5 * (f2 - f4) - (1.0 / 6.0) * r * U + (1.0 / 2.0) * r * V;
            p[7] = f5 + 0.5 * (f2 - f4) - (1.0 / 6.0) * r * U - (1.0 / 2.0) * r * V;
        }
    }
}

/* Zou-He velocity closures on a row.  src/bc/velocity_top.c:28-39 and
 * velocity_bottom.c:28-39.  The bottom variant stores its third closure into
 * slot 7 (velocity_bottom.c:39), so f6 keeps its streamed/stale value. */
static void zou_he_row(int Nx, int type, int j, const double *val1, const double *val2,
                       double *fn, const int *mask)
{
    for (int i = 0; i < Nx; i++) {
        if (mask[NODE(i, j)]) continue;
        double *p = &fn[POP(i, j, 0)];
        double U = val1[i], V = val2[i];
        double f0 = p[0], f1 = p[1], f2 = p[2], f3 = p[3], f4 = p[4];
        double f5 = p[5], f6 = p[6], f7 = p[7], f8 = p[8];
        if (type == ORC_BC_VELOCITY_TOP) {
            double r = (f0 + f1 + f3 + 2 * (f2 + f5 + f6)) / (1 + V);
            p[4] = f2 - 2.0 / 3.0 * r * V;
            p[7] = f5 + (f1 - f3) / 2.0 - 0.5 * r * U - 1.0 / 6.0 * r * V;
            p[8] = f6 - (f1 - f3) / 2.0 + 0.5 * r * U - 1.0 / 6.0 * r * V;
        } else {
            double r = (f0 + f1 + f3 + 2 * (f4 + f7 + f8)) / (1 - V);
            p[2] = f4 + 2.0 / 3.0 * r * V;
            p[5] = f7 - (f1 - f3) / 2.0 + 0.5 * r * U + 1.0 / 6.0 * r * V;
            p[7] = f8 + (f1 - f3) / 2.0 - 0.5 * r * U + 1.0 / 6.0 * r * V;
        }
    }
}

/* Regularized BC family, src/bc/..._regularized.c:18-71.  The boundary node gets
 * feq(rho_b, u_b) + 1.5*w*Q:Pi, where Pi is the second moment of
 * f_new - feq(lagged interior rho,u) at the neighbouring interior node.
 * Velocity variants take rho_b from the lagged interior density and u_b from
 * val1/val2; pressure variants take rho_b from val1 and u_b from the lagged
 * interior velocity. */
static void regularized(int Nx, int Ny, int on_column, int index, int inter, int pressure,
                        const double *val1, const double *val2, double *fn,
                        const double *rho, const double *u, const double *v, const int *mask)
{
    int n_edge = on_column ? Ny : Nx;
    for (int t = 0; t < n_edge; t++) {
        int ib = on_column ? index : t, jb = on_column ? t : index;
        int ii = on_column ? inter : t, ji = on_column ? t : inter;
        if (mask[NODE(ib, jb)]) continue;

        double ri = rho[NODE(ii, ji)], ui = u[NODE(ii, ji)], vi = v[NODE(ii, ji)];
        double usq_i = ui * ui + vi * vi;
        double rb, Ub, Vb;
        if (pressure) { rb = val1[t]; Ub = ui; Vb = vi; }
        else          { rb = ri; Ub = val1[t]; Vb = val2[t]; }
        double usq_b = Ub * Ub + Vb * Vb;

        double Pxx = 0.0, Pxy = 0.0, Pyy = 0.0;
        const double *pi = &fn[POP(ii, ji, 0)];
        for (int k = 0; k < Q; k++) {
            double neq = pi[k] - feq_c(k, ri, ui, vi, usq_i);
            Pxx += CX[k] * CX[k] * neq;
            Pxy += CX[k] * CY[k] * neq;
            Pyy += CY[k] * CY[k] * neq;
        }
        double *pb = &fn[POP(ib, jb, 0)];
        for (int k = 0; k < Q; k++) {
            double eq = feq_c(k, rb, Ub, Vb, usq_b);
            double Qxx = CX[k] * CX[k] - 1.0 / 3.0;
            double Qxy = CX[k] * CY[k];
            double Qyy = CY[k] * CY[k] - 1.0 / 3.0;
            double neq = (3.0 / 2.0) * W[k] * (Qxx * Pxx + 2.0 * Qxy * Pxy + Qyy * Pyy);
            pb[k] = eq + neq;
        }
    }
}

/* src/bc/convective_{inlet,outlet,top,bottom}.c:18-28 -- zero-gradient copy of
 * all nine f_new from the neighbouring column/row. */
static void convective(int Nx, int Ny, int on_column, int index, int from, double *fn,
                       const int *mask)
{
    int n_edge = on_column ? Ny : Nx;
    for (int t = 0; t < n_edge; t++) {
        int ib = on_column ? index : t, jb = on_column ? t : index;
        int is = on_column ? from : t, js = on_column ? t : from;
        if (mask[NODE(ib, jb)]) continue;
        for (int k = 0; k < Q; k++) fn[POP(ib, jb, k)] = fn[POP(is, js, k)];
    }
}

/* src/bc/periodic_x.c:19-50 and periodic_y.c:19-50 -- re-pull the wrapped
 * populations of the first/last column (row) from post-collision f; the other
 * coordinate is not wrapped. */
static void periodic(int Nx, int Ny, int along_x, const double *f, double *fn,
                     const int *mask)
{
    for (int k = 0; k < Q; k++) {
        int c = along_x ? CX[k] : CY[k];
        if (c == 0) continue;
        int n_edge = along_x ? Ny : Nx;
        int dst = c > 0 ? 0 : (along_x ? Nx : Ny) - 1;
        int src = c > 0 ? (along_x ? Nx : Ny) - c : -c - 1;
        for (int t = 0; t < n_edge; t++) {
            int i = along_x ? dst : t, j = along_x ? t : dst;
            int is = along_x ? src : t - CX[k], js = along_x ? t - CY[k] : src;
            if (along_x ? (js < 0 || js >= Ny) : (is < 0 || is >= Nx)) continue;
            if (mask[NODE(i, j)]) continue;
            if (mask[NODE(is, js)]) continue;
            fn[POP(i, j, k)] = f[POP(is, js, k)];
        }
    }
}

void orc_apply_bc(int Nx, int Ny, const orc_bc *bc, const double *f, double *f_new,
                  const double *rho, const double *u, const double *v, const int *mask)
{
    int x = bc->index;
    switch (bc->type) {
    case ORC_BC_VELOCITY_INLET:
    case ORC_BC_PRESSURE_INLET:
    case ORC_BC_VELOCITY_OUTLET:
    case ORC_BC_PRESSURE_OUTLET:
        zou_he_column(Nx, Ny, bc->type, x, bc->val1, bc->val2, f_new, v, mask);
        break;
    case ORC_BC_VELOCITY_TOP:
    case ORC_BC_VELOCITY_BOTTOM:
        zou_he_row(Nx, bc->type, x, bc->val1, bc->val2, f_new, mask);
        break;
    case ORC_BC_VELOCITY_INLET_REG:
        regularized(Nx, Ny, 1, x, x + 1, 0, bc->val1, bc->val2, f_new, rho, u, v, mask); break;
    case ORC_BC_VELOCITY_OUTLET_REG:
        regularized(Nx, Ny, 1, x, x - 1, 0, bc->val1, bc->val2, f_new, rho, u, v, mask); break;
    case ORC_BC_VELOCITY_TOP_REG:
        regularized(Nx, Ny, 0, x, x - 1, 0, bc->val1, bc->val2, f_new, rho, u, v, mask); break;
    case ORC_BC_VELOCITY_BOTTOM_REG:
        regularized(Nx, Ny, 0, x, x + 1, 0, bc->val1, bc->val2, f_new, rho, u, v, mask); break;
    case ORC_BC_PRESSURE_INLET_REG:
        regularized(Nx, Ny, 1, x, x + 1, 1, bc->val1, bc->val2, f_new, rho, u, v, mask); break;
    case ORC_BC_PRESSURE_OUTLET_REG:
        regularized(Nx, Ny, 1, x, x - 1, 1, bc->val1, bc->val2, f_new, rho, u, v, mask); break;
    case ORC_BC_PRESSURE_TOP_REG:
        regularized(Nx, Ny, 0, x, x - 1, 1, bc->val1, bc->val2, f_new, rho, u, v, mask); break;
    case ORC_BC_PRESSURE_BOTTOM_REG:
        regularized(Nx, Ny, 0, x, x + 1, 1, bc->val1, bc->val2, f_new, rho, u, v, mask); break;
    case ORC_BC_CONVECTIVE_INLET:  convective(Nx, Ny, 1, x, x + 1, f_new, mask); break;
    case ORC_BC_CONVECTIVE_OUTLET: convective(Nx, Ny, 1, x, x - 1, f_new, mask); break;
    case ORC_BC_CONVECTIVE_TOP:    convective(Nx, Ny, 0, x, x - 1, f_new, mask); break;
    case ORC_BC_CONVECTIVE_BOTTOM: convective(Nx, Ny, 0, x, x + 1, f_new, mask); break;
    case ORC_BC_PERIODIC_X: periodic(Nx, Ny, 1, f, f_new, mask); break;
    case ORC_BC_PERIODIC_Y: periodic(Nx, Ny, 0, f, f_new, mask); break;
    default: break;
    }
}

/* src/core/main_lbm.c:31-61 -- collide, BB/IBB, force, stream, BC table in
 * order, moments.  The force pointers are passed as (D, L) like main_lbm.c:43. */
void orc_step(int Nx, int Ny, double *f, double *f_new, double *rho, double *u, double *v,
              const double *omega, const int *mask, const orc_bc *bcs, int nbc,
              int per_x, int per_y, double Fx, double Fy, double *L, double *D,
              const double *phi, int use_ibb)
{
    orc_collision(Nx, Ny, f, rho, u, v, omega, mask, Fx, Fy);
    if (use_ibb) orc_interpolated_bounce_back(Nx, Ny, f, f_new, mask, phi, per_x, per_y);
    else orc_bounce_back(Nx, Ny, f, f_new, mask, per_x, per_y);
    orc_force_mem(Nx, Ny, f, f_new, mask, per_x, per_y, D, L);
    orc_streaming(Nx, Ny, f, f_new, mask, per_x, per_y);
    for (int b = 0; b < nbc; b++) orc_apply_bc(Nx, Ny, &bcs[b], f, f_new, rho, u, v, mask);
    orc_moments(Nx, Ny, f_new, mask, rho, u, v, Fx, Fy);
}

void orc_run(int Nx, int Ny, int nsteps, double *buf_a, double *buf_b, double *rho,
             double *u, double *v, const double *omega, const int *mask,
             const orc_bc *bcs, int nbc, int per_x, int per_y, double Fx, double Fy,
             double *L, double *D, const double *phi, int use_ibb, int *swapped)
{
    double *f = buf_a, *fn = buf_b;
    for (int it = 0; it < nsteps; it++) {
        orc_step(Nx, Ny, f, fn, rho, u, v, omega, mask, bcs, nbc, per_x, per_y, Fx, Fy,
                 &L[it], &D[it], phi, use_ibb);
        double *t = f; f = fn; fn = t;
    }
    *swapped = (f != buf_a);
}

/* src/core/mei_initialization.c:33-139 -- Mei et al. consistent initialization:
 * incompressible equilibrium around the mean density, then repeated
 * {rho = sum f; collide with u held fixed (no solid test); stream; swap} until
 * max|drho| <= tol.  The swap is local, so after an odd number of iterations
 * the caller's f holds the last post-collision field and f_new the streamed one. */
int orc_mei(int Nx, int Ny, double *f, double *f_new, const double *rho_in,
            const double *U_in, const double *V_in, double *rho, double *u, double *v,
            const double *omega, int per_x, int per_y, int iter_max, double tol,
            const int *mask, double *max_drho_out)
{
    size_t nn = (size_t)Nx * (size_t)Ny;
    memcpy(u, U_in, nn * sizeof(double));
    memcpy(v, V_in, nn * sizeof(double));
    memcpy(rho, rho_in, nn * sizeof(double));

    double rho0 = 0.0;
    for (size_t n = 0; n < nn; n++) rho0 += rho[n];
    rho0 /= (Nx * Ny);

    #pragma omp parallel for
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) {
            size_t n = NODE(i, j);
            double ux = u[n], uy = v[n];
            double usq = ux * ux + uy * uy;
            for (int k = 0; k < Q; k++) {
                double cu = CX[k] * ux + CY[k] * uy;
                f[n * Q + k] = W[k] * rho[n] + W[k] * rho0 * (3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
            }
        }
    }

    double *prev = malloc(nn * sizeof(double));
    double max_d = 0.0;
    int it, ran = 0;
    for (it = 0; it < iter_max; it++) {
        ran = it + 1;
        memcpy(prev, rho, nn * sizeof(double));
        #pragma omp parallel for
        for (int i = 0; i < Nx; i++) {
            for (int j = 0; j < Ny; j++) {
                size_t n = NODE(i, j);
                double s = 0.0;
                for (int k = 0; k < Q; k++) s += f[n * Q + k];
                rho[n] = s;
            }
        }
        #pragma omp parallel for
        for (int i = 0; i < Nx; i++) {
            for (int j = 0; j < Ny; j++) {
                size_t n = NODE(i, j);
                double ux = u[n], uy = v[n];
                double usq = ux * ux + uy * uy;
                double om = omega[n];
                for (int k = 0; k < Q; k++) {
                    double cu = CX[k] * ux + CY[k] * uy;
                    double eq = W[k] * rho[n] + W[k] * rho0 * (3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
                    f[n * Q + k] += -om * (f[n * Q + k] - eq);
                }
            }
        }
        orc_streaming(Nx, Ny, f, f_new, mask, per_x, per_y);
        double *t = f; f = f_new; f_new = t;

        max_d = 0.0;
        for (size_t n = 0; n < nn; n++) {
            double d = fabs(rho[n] - prev[n]);
            if (d > max_d) max_d = d;
        }
        if (max_d <= tol) break;
    }
    free(prev);
    if (max_drho_out) *max_drho_out = max_d;
    return ran;
}

/* src/core/sponge_layer.c:21-44 -- radial relaxation-time ramp. */
void orc_sponge_layer(int Nx, int Ny, double dx, double xc, double yc, double r_inner,
                      double r_outer, double tau_base, double incr_tau, double exp_coeff,
                      double *omega)
{
    for (int j = 0; j < Ny; j++) {
        double y = j * dx - yc;
        for (int i = 0; i < Nx; i++) {
            double x = i * dx - xc;
            double r = sqrt(x * x + y * y);
            double tau;
            if (r <= r_inner) tau = tau_base;
            else if (r <= r_outer)
                tau = tau_base + (incr_tau - 1) * tau_base * (exp(exp_coeff * (r - r_inner)) - 1)
                                 / (exp(exp_coeff * (r_outer - r_inner)) - 1);
            else tau = tau_base * incr_tau;
            omega[NODE(i, j)] = 1.0 / tau;
        }
    }
}

/* cases/nit/nit.c:13-47.  The reference calls its Euler update with the output
 * arrays aliasing the inputs, so the chi update already sees the new psi. */
double orc_tdibc_velocity(orc_tdibc *s, double p, double dt)
{
    for (int i = 0; i < 2; i++) {
        s->phi[i] = s->phi[i] + dt * (p + s->lambdas[i] * s->phi[i]);
        s->psi[i] = s->psi[i] + dt * (p + s->alpha[i] * s->psi[i] - s->beta[i] * s->chi[i]);
        s->chi[i] = s->chi[i] + dt * (s->beta[i] * s->psi[i] + s->alpha[i] * s->chi[i]);
    }
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = 0; i < 2; i++) {
        a += s->A[i] * s->phi[i];
        b += s->B[i] * s->psi[i];
        c += s->C[i] * s->chi[i];
    }
    return s->Zinf * p - a - 2.0 * (b + c);
}

/* cases/nit/nit.c:90-130 -- physical-to-lattice conversion of the NIT case. */
void orc_nit_config(double p_a, double f_exc, orc_nit_cfg *c)
{
    (void)p_a;
    double c0 = 340.0, nu = 1.5e-5;
    int tot_cycles = 1 + 10 + 2;
    c->n_orifices = 2;
    c->variant = 0;
    c->Ly = 0.009906; c->y1 = 0.002528; c->y2 = 0.007478; c->d = 0.00117;
    c->tau = 0.5005; c->rho0 = 1.184;
    c->cs = 1.0 / sqrt(3.0);
    c->dx = 3 * nu * c->cs / (c0 * (c->tau - 0.5));
    c->dt = c->dx * c->dx * (c->tau - 0.5) / (3 * nu);
    c->cu = c->dx / c->dt;
    c->cp = 1.0 * c->cu * c->cu;
    double lambda = c0 / f_exc;
    double Lx = tot_cycles * lambda;
    c->Nx = (int)ceil(Lx / c->dx);
    c->Ny = (int)ceil(c->Ly / c->dx);
    c->nt = (int)ceil(tot_cycles / (f_exc * c->dt));
    c->init_iter_max = 100;
    c->tol_rho = 1e-10;
}

/* cases/nit/nit.c:120-311 -- initial packet, Mei, then the TDIBC-driven loop. */
int orc_nit_run(const orc_nit_cfg *c, double p_a, double f_exc, orc_tdibc *imp, int nt,
                double *rec, double *f_out, double *rho_out)
{
    int Nx = c->Nx, Ny = c->Ny;
    size_t nn = (size_t)Nx * Ny;
    double c0 = 340.0;
    double lambda = c0 / f_exc;
    double Lx = (1 + 10 + 2) * lambda;
    double kx = 2 * M_PI / lambda;
    double rho_a = (p_a / c->cp) / (c->cs * c->cs);
    double x_end = Lx - 1 * lambda;
    double x_start = x_end - 10 * lambda;

    double *rho_in = malloc(nn * 8), *U_in = malloc(nn * 8), *V_in = malloc(nn * 8);
    double *rho = malloc(nn * 8), *u = malloc(nn * 8), *v = malloc(nn * 8);
    double *omega = malloc(nn * 8), *phi = calloc(nn, 8);
    int *mask = calloc(nn, sizeof(int));
    double *fa = calloc(nn * Q, 8), *fb = calloc(nn * Q, 8);
    for (size_t n = 0; n < nn; n++) omega[n] = 1.0 / c->tau;
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) {
            double x = i * c->dx;
            double env = ((x >= x_start) & (x <= x_end)) ? 1.0 : 0.0;
            double rp = rho_a * env * sin(kx * x);
            rho_in[NODE(i, j)] = c->rho0 + rp;
            U_in[NODE(i, j)] = c->cs * rp / c->rho0;
            V_in[NODE(i, j)] = 0.0;
        }
    }
    int iters = orc_mei(Nx, Ny, fa, fb, rho_in, U_in, V_in, rho, u, v, omega, 0, 1,
                        c->init_iter_max, c->tol_rho, mask, NULL);

    double *r0 = malloc(Ny * 8), *zero = calloc(Ny, 8), *vbc = malloc(Ny * 8);
    for (int j = 0; j < Ny; j++) r0[j] = c->rho0;
    orc_bc bcs[3] = {
        {ORC_BC_PERIODIC_Y, 0, NULL, NULL},
        {ORC_BC_PRESSURE_INLET, 0, r0, NULL},
        {ORC_BC_VELOCITY_OUTLET_REG, Nx - 1, vbc, zero},
    };
    double p = 0.0, vac = 0.0, L = 0.0, D = 0.0;
    double *f = fa, *fn = fb;
    for (int it = 0; it < nt; it++) {
        rec[3 * it + 0] = it * c->dt;
        rec[3 * it + 1] = p;
        rec[3 * it + 2] = vac;
        double rs = 0.0;
        for (int j = 0; j < Ny; j++) rs += rho[NODE(Nx - 1, j)];
        rs /= Ny;
        p = c->cs * c->cs * (rs - c->rho0) * c->cp;
        vac = orc_tdibc_velocity(imp, p, c->dt);
        double tv = (4 * c->Ly * c->Ly) / (c->n_orifices * c->d * c->d) * (vac / c->cu);
        for (int j = 0; j < Ny; j++) {
            double y = j * c->dx;
            int hole = (y >= c->y1 - c->d / 2 && y <= c->y1 + c->d / 2)
                     | (y >= c->y2 - c->d / 2 && y <= c->y2 + c->d / 2);
            vbc[j] = hole ? tv : 0.0;
            if (c->variant == 1) vbc[j] = vac / c->cu;
        }
        orc_step(Nx, Ny, f, fn, rho, u, v, omega, mask, bcs, 3, 0, 1, 0.0, 0.0, &L, &D,
                 phi, 0);
        double *t = f; f = fn; fn = t;
    }
    if (f_out) memcpy(f_out, f, nn * Q * 8);
    if (rho_out) memcpy(rho_out, rho, nn * 8);
    free(rho_in); free(U_in); free(V_in); free(rho); free(u); free(v); free(omega);
    free(phi); free(mask); free(fa); free(fb); free(r0); free(zero); free(vbc);
    return iters;
}

/* ---------------------------------------------------------------------------
 * Acoustic-liner duct (see lbm_oracle.h).  Composes the restated reference
 * routines: orc_step with the table {convective_inlet(0), convective_outlet(Nx-1),
 * velocity_bottom_regularized(0)} and one orc_tdibc per liner node.
 * ------------------------------------------------------------------------- */
void orc_liner_setup(const orc_liner_cfg *c, double *f, double *rho, double *u, double *v,
                     double *omega, int *mask)
{
    int Nx = c->Nx, Ny = c->Ny;
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) {
            size_t n = NODE(i, j);
            double xr = (i - c->x0) / c->halfwidth;
            rho[n] = c->rho0 + c->amp * exp(-log(2.0) * xr * xr);
            u[n] = 0.0;
            v[n] = 0.0;
            int wall = (j == Ny - 1) || (j == 0 && (i < c->liner_x0 || i > c->liner_x1));
            mask[n] = wall;
            double tau = c->tau;
            int d = i < Nx - 1 - i ? i : Nx - 1 - i;   /* distance to the nearer end */
            if (c->sponge > 0 && d < c->sponge)
                tau = c->tau_sponge + (c->tau - c->tau_sponge) * ((double)d / c->sponge);
            omega[n] = 1.0 / tau;
            for (int k = 0; k < Q; k++) f[n * Q + k] = feq_c(k, rho[n], 0.0, 0.0, 0.0);
        }
    }
}

void orc_liner_run(const orc_liner_cfg *c, int nt, double *probe_hist, double *wall_p,
                   double *f_out, double *rho_out)
{
    int Nx = c->Nx, Ny = c->Ny;
    size_t nn = (size_t)Nx * Ny;
    double *fa = malloc(nn * Q * 8), *fb = malloc(nn * Q * 8);
    double *rho = malloc(nn * 8), *u = malloc(nn * 8), *v = malloc(nn * 8);
    double *omega = malloc(nn * 8), *phi = calloc(nn, 8);
    int *mask = malloc(nn * sizeof(int));
    orc_liner_setup(c, fa, rho, u, v, omega, mask);
    memcpy(fb, fa, nn * Q * 8);

    double *zero = calloc(Ny > Nx ? Ny : Nx, 8), *vbc = calloc(Nx, 8);
    int nl = c->liner_x1 - c->liner_x0 + 1;
    orc_tdibc *st = malloc(nl * sizeof(orc_tdibc));
    for (int l = 0; l < nl; l++) st[l] = c->imp;
    orc_bc bcs[3] = {
        {ORC_BC_CONVECTIVE_INLET, 0, NULL, NULL},
        {ORC_BC_CONVECTIVE_OUTLET, Nx - 1, NULL, NULL},
        {ORC_BC_VELOCITY_BOTTOM_REG, 0, zero, vbc},
    };
    const double cs = 1.0 / sqrt(3.0);
    double L = 0.0, D = 0.0;
    double *f = fa, *fn = fb;
    for (int it = 0; it < nt; it++) {
        if (probe_hist) probe_hist[it] = rho[NODE(c->probe_x, c->probe_y)];
        for (int l = 0; l < nl; l++) {
            int i = c->liner_x0 + l;
            double p = cs * cs * (rho[NODE(i, 0)] - c->rho0) * 1.0;
            double va = orc_tdibc_velocity(&st[l], p, 1.0);
            vbc[i] = c->gain * (va / 1.0);
        }
        orc_step(Nx, Ny, f, fn, rho, u, v, omega, mask, bcs, 3, 0, 0, 0.0, 0.0, &L, &D, phi, 0);
        double *t = f; f = fn; fn = t;
    }
    if (wall_p)
        for (int l = 0; l < nl; l++)
            wall_p[l] = cs * cs * (rho[NODE(c->liner_x0 + l, 0)] - c->rho0) * 1.0;
    if (f_out) memcpy(f_out, f, nn * Q * 8);
    if (rho_out) memcpy(rho_out, rho, nn * 8);
    free(fa); free(fb); free(rho); free(u); free(v); free(omega); free(phi); free(mask);
    free(zero); free(vbc); free(st);
}
