/*
 * lbm_utils.h -- output helpers of the lbc C API (host I/O).
 * Prototypes identical to the reference's include/lbm_utils.h:4-16.
 */
#ifndef LBM_UTILS_H
#define LBM_UTILS_H

int ensure_directory_exists(const char *path);

void write_vtk_binary_2D(
    const char *filename,
    int Nx,
    int Ny,
    double dx,
    double u[Ny][Nx],
    double v[Ny][Nx],
    double rho[Ny][Nx],
    double cu,
    double crho
);

#endif
