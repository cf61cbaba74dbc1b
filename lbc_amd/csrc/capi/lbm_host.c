/*
 * lbm_host.c -- host-side set-up and output routines of the lbc C API.
 *
 * These are not part of the time step: sponge_layer fills omega_eff once before
 * a run, and the two utilities write files.  Formats and messages are kept
 * compatible with the reference so its post-processing scripts read the output.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include "lbm.h"

/* src/utils/filesystem.c:38-57 -- 0 if `path` is (or now is) a directory. */
int ensure_directory_exists(const char *path)
{
    struct stat st;
    if (stat(path, &st) == 0) {
        if (S_ISDIR(st.st_mode)) return 0;
        fprintf(stderr, "Path exists but is not a directory: %s\n", path);
        return -1;
    }
    if (mkdir(path, 0755) != 0) {
        perror("mkdir failed");
        return -1;
    }
    return 0;
}

static uint32_t to_big_endian_bits(float x)
{
    uint32_t b;
    memcpy(&b, &x, 4);
#if __BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__
    b = __builtin_bswap32(b);
#endif
    return b;
}

/* The text parts of the VTK file, shared with the device-packed writer
 * (lbmx_write_vtk): everything before the velocity block, and between the
 * velocity and density blocks. */
__attribute__((visibility("hidden"))) void lbm_vtk_head(FILE *fd, int Nx, int Ny, double dx)
{
    fprintf(fd, "# vtk DataFile Version 3.0\n2D LBM output\nBINARY\n");
    fprintf(fd, "DATASET STRUCTURED_POINTS\nDIMENSIONS %d %d 1\nORIGIN 0 0 0\n", Nx, Ny);
    fprintf(fd, "SPACING %f %f 1\n\nPOINT_DATA %d\n", dx, dx, Nx * Ny);
    fprintf(fd, "VECTORS velocity float\n");
}

__attribute__((visibility("hidden"))) void lbm_vtk_mid(FILE *fd)
{
    fprintf(fd, "\nSCALARS density float 1\nLOOKUP_TABLE default\n");
}

/* Legacy VTK STRUCTURED_POINTS file with big-endian float32 VECTORS velocity and
 * SCALARS density, the layout src/utils/write_vtk_binary_2D.c:46-78 produces.
 * Unlike the reference this version keeps no static buffer (re-entrant). */
void write_vtk_binary_2D(const char *filename, int Nx, int Ny, double dx, double u[Ny][Nx],
                         double v[Ny][Nx], double rho[Ny][Nx], double cu, double crho)
{
    FILE *fd = fopen(filename, "wb");
    if (!fd) {
        perror("Failed to open VTK file");
        return;
    }
    lbm_vtk_head(fd, Nx, Ny, dx);

    const size_t nn = (size_t)Nx * (size_t)Ny;
    uint32_t *vel = malloc(nn * 3 * sizeof(uint32_t));
    uint32_t *den = malloc(nn * sizeof(uint32_t));
    if (!vel || !den) {
        fprintf(stderr, "write_vtk_binary_2D: out of memory\n");
        free(vel);
        free(den);
        fclose(fd);
        return;
    }
    const uint32_t zero = to_big_endian_bits(0.0f);
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) {
            const size_t n = (size_t)j * Nx + i;
            vel[3 * n + 0] = to_big_endian_bits((float)(u[j][i] * cu));
            vel[3 * n + 1] = to_big_endian_bits((float)(v[j][i] * cu));
            vel[3 * n + 2] = zero;
            den[n] = to_big_endian_bits((float)(rho[j][i] * crho));
        }
    }
    fwrite(vel, sizeof(uint32_t), nn * 3, fd);
    lbm_vtk_mid(fd);
    fwrite(den, sizeof(uint32_t), nn, fd);
    free(vel);
    free(den);
    fclose(fd);
}

/* src/core/sponge_layer.c:21-44 -- radial relaxation-time profile (set-up):
 * tau = tau_base inside r_inner, tau_base*incr_tau outside r_outer, and an
 * exponential ramp between; omega_eff = 1/tau. */
void sponge_layer(int Nx, int Ny, double dx, double xc, double yc, double r_inner,
                  double r_outer, double tau_base, double incr_tau, double exp_coeff,
                  double omega_eff[Ny][Nx])
{
    const double ramp_den = exp(exp_coeff * (r_outer - r_inner)) - 1;
    for (int j = 0; j < Ny; j++) {
        const double y = j * dx - yc;
        for (int i = 0; i < Nx; i++) {
            const double x = i * dx - xc;
            const double r = sqrt(x * x + y * y);
            double tau;
            if (r <= r_inner)
                tau = tau_base;
            else if (r <= r_outer)
                tau = tau_base + (incr_tau - 1) * tau_base * (exp(exp_coeff * (r - r_inner)) - 1) /
                                     ramp_den;
            else
                tau = tau_base * incr_tau;
            omega_eff[j][i] = 1.0 / tau;
        }
    }
}
