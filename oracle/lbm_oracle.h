/*
 * lbm_oracle.h -- CPU restatement of the lbc D2Q9 BGK step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load it.  Nothing in the product links or calls it.
 *
 * Layout is the reference's: populations are array-of-structures
 * f[(j*Nx + i)*9 + k], fields are row-major a[j*Nx + i], solid_mask is int.
 * Every routine follows the reference's loop order and floating-point operation
 * order so that, compiled without FMA contraction, results are bit-identical to
 * the reference build (pinned in tests/test_oracle_vs_reference.py).
 *
 * Lattice constants are the reference driver's D2Q9 tables
 * (cases/nit/nit.c:195-202).
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary-condition codes.  The numbering matches LBMX_BC_* in include/lbmx.h. */
enum {
    ORC_BC_NONE = 0,
    ORC_BC_CONVECTIVE_INLET = 1,
    ORC_BC_CONVECTIVE_OUTLET = 2,
    ORC_BC_PRESSURE_INLET = 3,
    ORC_BC_PRESSURE_OUTLET = 4,
    ORC_BC_PERIODIC_X = 5,
    ORC_BC_PERIODIC_Y = 6,
    ORC_BC_VELOCITY_INLET = 7,
    ORC_BC_VELOCITY_OUTLET = 8,
    ORC_BC_VELOCITY_TOP = 9,
    ORC_BC_VELOCITY_BOTTOM = 10,
    ORC_BC_VELOCITY_OUTLET_REG = 11,
    ORC_BC_VELOCITY_INLET_REG = 12,
    ORC_BC_VELOCITY_TOP_REG = 13,
    ORC_BC_VELOCITY_BOTTOM_REG = 14,
    ORC_BC_PRESSURE_OUTLET_REG = 15,
    ORC_BC_PRESSURE_INLET_REG = 16,
    ORC_BC_PRESSURE_TOP_REG = 17,
    ORC_BC_PRESSURE_BOTTOM_REG = 18,
    ORC_BC_CONVECTIVE_TOP = 19,
    ORC_BC_CONVECTIVE_BOTTOM = 20
};

typedef struct {
    int type;            /* ORC_BC_* */
    int index;           /* column (inlet/outlet) or row (top/bottom) */
    const double *val1;  /* per-node values, may be NULL when unused */
    const double *val2;
} orc_bc;

/* src/core/collision.c:1-41 */
void orc_collision(int Nx, int Ny, double *f, const double *rho, const double *u,
                   const double *v, const double *omega, const int *mask,
                   double Fx, double Fy);
/* src/core/streaming.c:1-45 */
void orc_streaming(int Nx, int Ny, const double *f, double *f_new, const int *mask,
                   int per_x, int per_y);
/* src/bc/bounce_back.c:1-50 */
void orc_bounce_back(int Nx, int Ny, const double *f, double *f_new, const int *mask,
                     int per_x, int per_y);
/* src/bc/interpolated_bounce_back.c:4-104 */
void orc_interpolated_bounce_back(int Nx, int Ny, const double *f, double *f_new,
                                  const int *mask, const double *phi, int per_x, int per_y);
/* src/core/compute_force_mem.c:1-72 (serial, row-major summation order) */
void orc_force_mem(int Nx, int Ny, const double *f, const double *f_new, const int *mask,
                   int per_x, int per_y, double *Fx_out, double *Fy_out);
/* src/core/compute_macroscopic_fields.c:4-49 */
void orc_moments(int Nx, int Ny, const double *f_new, const int *mask, double *rho,
                 double *u, double *v, double Fx, double Fy);
/* One boundary condition from src/bc/ files, dispatched on bc->type. */
void orc_apply_bc(int Nx, int Ny, const orc_bc *bc, const double *f, double *f_new,
                  const double *rho, const double *u, const double *v, const int *mask);
/* src/core/main_lbm.c:5-62 */
void orc_step(int Nx, int Ny, double *f, double *f_new, double *rho, double *u, double *v,
              const double *omega, const int *mask, const orc_bc *bcs, int nbc,
              int per_x, int per_y, double Fx, double Fy, double *L, double *D,
              const double *phi, int use_ibb);
/* Run nsteps of orc_step with the driver pointer swap (cases/nit/nit.c:306-311).
 * On return buf_a/buf_b hold what the driver's f/f_new point to after the loop
 * when nsteps is even; for odd nsteps the roles are exchanged (*swapped = 1). */
void orc_run(int Nx, int Ny, int nsteps, double *buf_a, double *buf_b, double *rho,
             double *u, double *v, const double *omega, const int *mask,
             const orc_bc *bcs, int nbc, int per_x, int per_y, double Fx, double Fy,
             double *L, double *D, const double *phi, int use_ibb, int *swapped);
/* src/core/mei_initialization.c:7-142.  Returns the number of iterations run. */
int orc_mei(int Nx, int Ny, double *f, double *f_new, const double *rho_in,
            const double *U_in, const double *V_in, double *rho, double *u, double *v,
            const double *omega, int per_x, int per_y, int iter_max, double tol,
            const int *mask, double *max_drho_out);
/* src/core/sponge_layer.c:3-46 */
void orc_sponge_layer(int Nx, int Ny, double dx, double xc, double yc, double r_inner,
                      double r_outer, double tau_base, double incr_tau, double exp_coeff,
                      double *omega);

/* Time-domain impedance BC of cases/nit/nit.c:13-47: explicit-Euler ADE update of
 * the N=2 real poles and N=2 complex pole pairs, then the wall velocity. */
typedef struct {
    double Zinf, A[2], lambdas[2], B[2], C[2], alpha[2], beta[2];
    double phi[2], psi[2], chi[2];
} orc_tdibc;
double orc_tdibc_velocity(orc_tdibc *s, double p_fluc, double dt);

/* The nit impedance-liner driver loop (cases/nit/nit.c:77-320) without file
 * output: Mei initialization, then nt steps.  rec receives nt rows of
 * (t, p_fluc, v_ac) exactly as nit.c:273 prints them (values lagged one step). */
typedef struct {
    int Nx, Ny, nt, init_iter_max;
    double dx, dt, cu, cp, cs, rho0, tau, Ly, y1, y2, d, tol_rho;
    int n_orifices;
    /* 0: orifice rows get the area-scaled velocity (nit.c:286-300, the current
     * source); 1: every row gets v_ac/cu ("uniform" run of nit/process.py:428) */
    int variant;
} orc_nit_cfg;
void orc_nit_config(double p_a, double f_exc, orc_nit_cfg *cfg);
int orc_nit_run(const orc_nit_cfg *cfg, double p_a, double f_exc, orc_tdibc *imp, int nt,
                double *rec, double *f_out, double *rho_out);

/* Acoustic-liner duct driver (the impedance-liner configuration built from the
 * reference's components): top wall solid, bottom wall solid outside
 * [liner_x0, liner_x1] where row 0 is a velocity_bottom_regularized boundary
 * whose normal velocity comes from one TDIBC ODE per node (orc_tdibc_velocity,
 * nit.c:13-47) driven by that node's lagged density; duct ends are convective
 * (zero-gradient) columns behind a sponge in which tau ramps linearly from tau
 * to tau_sponge over `sponge` cells.  Initial state: Gaussian density pulse
 * rho = rho0 + amp*exp(-ln2 ((x-x0)/halfwidth)^2) at rest, populations at
 * equilibrium in both buffers.
 * probe_hist receives the lagged density at (probe_x, probe_y) for each step;
 * wall_p the lagged p' at every liner node after the last step. */
typedef struct {
    int Nx, Ny, liner_x0, liner_x1, sponge, probe_x, probe_y;
    double tau, tau_sponge, rho0, amp, x0, halfwidth, gain;
    orc_tdibc imp;   /* coefficients; ODE state starts at zero */
} orc_liner_cfg;
void orc_liner_setup(const orc_liner_cfg *c, double *f, double *rho, double *u, double *v,
                     double *omega, int *mask);
void orc_liner_run(const orc_liner_cfg *c, int nt, double *probe_hist, double *wall_p,
                   double *f_out, double *rho_out);

#ifdef __cplusplus
}
#endif
#endif
