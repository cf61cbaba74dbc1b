/*
 * lbm_bc.h -- boundary conditions of the lbc C API, served by the MI355X library.
 *
 * After preprocessing, every prototype is identical to the reference's
 * include/lbm_bc.h:4-355.  The eighteen functions with the BCFunction signature
 * are the plug-ins of the Boundary table: main_lbm recognises them by address and
 * runs them as device kernels in table order; called directly, each uploads its
 * arrays, runs on the device and writes f_new back.
 *
 * convective_top / convective_bottom exist in the reference as definitions
 * without a declaration (src/bc/convective_top.c, convective_bottom.c); they are
 * declared here as well.
 */
#ifndef LBM_BC_H
#define LBM_BC_H

#include "lbm_types.h"

/* The parameter list shared by every BCFunction (lbm_types.h). */
#define LBM_BC_PARAMETERS                                                          \
    int Nx, int Ny, int Q, int index, const double *val1, const double *val2,     \
    double f[Ny][Nx][Q], double f_new[Ny][Nx][Q], double rho[Ny][Nx],             \
    double u[Ny][Nx], double v[Ny][Nx], int solid_mask[Ny][Nx], int cx[],         \
    int cy[], double w[]

void bounce_back(
    int Nx,
    int Ny,
    int Q,
    double f[Ny][Nx][Q],
    double f_new[Ny][Nx][Q],
    int solid_mask[Ny][Nx],
    int cx[],
    int cy[],
    int opp[],
    int isperiodic_x,
    int isperiodic_y
);

void interpolated_bounce_back(
    int Nx,
    int Ny,
    int Q,
    double f[Ny][Nx][Q],
    double f_new[Ny][Nx][Q],
    int solid_mask[Ny][Nx],
    double phi[Ny][Nx],
    int opp[],
    int cx[],
    int cy[],
    int isperiodic_x,
    int isperiodic_y
);

void convective_inlet(LBM_BC_PARAMETERS);
void convective_outlet(LBM_BC_PARAMETERS);
void pressure_inlet(LBM_BC_PARAMETERS);
void pressure_outlet(LBM_BC_PARAMETERS);
void periodic_x(LBM_BC_PARAMETERS);
void periodic_y(LBM_BC_PARAMETERS);
void velocity_inlet(LBM_BC_PARAMETERS);
void velocity_outlet(LBM_BC_PARAMETERS);
void velocity_top(LBM_BC_PARAMETERS);
void velocity_bottom(LBM_BC_PARAMETERS);
void velocity_outlet_regularized(LBM_BC_PARAMETERS);
void velocity_inlet_regularized(LBM_BC_PARAMETERS);
void velocity_top_regularized(LBM_BC_PARAMETERS);
void velocity_bottom_regularized(LBM_BC_PARAMETERS);
void pressure_outlet_regularized(LBM_BC_PARAMETERS);
void pressure_inlet_regularized(LBM_BC_PARAMETERS);
void pressure_top_regularized(LBM_BC_PARAMETERS);
void pressure_bottom_regularized(LBM_BC_PARAMETERS);

/* defined but undeclared in the reference */
void convective_top(LBM_BC_PARAMETERS);
void convective_bottom(LBM_BC_PARAMETERS);

#endif
