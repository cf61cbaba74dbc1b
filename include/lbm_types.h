/*
 * lbm_types.h -- boundary-condition plug-in types of the lbc C API.
 *
 * Declarations are identical to the reference's include/lbm_types.h:4-27 so that
 * its drivers compile and link against this library unchanged.  A boundary is
 * recognised by the address of its apply function: the drivers themselves test
 * `apply == periodic_x` (cases/nit/nit.c:189-192), and main_lbm maps this
 * library's own BC functions onto device kernels the same way.
 */
#ifndef LBM_TYPES_H
#define LBM_TYPES_H

typedef void (*BCFunction)(
    int Nx,
    int Ny,
    int Q,
    int index,
    const double *val1,
    const double *val2,
    double f[Ny][Nx][Q],
    double f_new[Ny][Nx][Q],
    double rho[Ny][Nx],
    double u[Ny][Nx],
    double v[Ny][Nx],
    int solid_mask[Ny][Nx],
    int cx[],
    int cy[],
    double w[]
);

typedef struct {
    int index;
    const double *val1;
    const double *val2;
    BCFunction apply;
} Boundary;

#endif
