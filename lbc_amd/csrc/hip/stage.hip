// stage.hip -- device implementations of the reference's stand-alone functions
// (lbm_core.h:33-115, lbm_bc.h:4-355) and of Mei's consistent initialization.
//
// Each stage takes the caller's host arrays, runs on a device lattice cached per
// lattice size and writes its outputs back; the arithmetic is the fused step's
// (same kernels or the same device functions).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "lattice_impl.h"
#include "lbmx.h"
#include "step_kernels.h"

using namespace lbmx;

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// src/core/streaming.c:14-44 -- pull into fluid nodes, skipping out-of-domain
// and solid sources (those entries of D are left as they were).
__global__ void __launch_bounds__(256) k_stream_pull(const double *__restrict__ S,
                                                     double *__restrict__ D, Geo g,
                                                     const int *__restrict__ solid)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= g.Nx) return;
    const size_t n = (size_t)y * g.pitch + x;
    if (solid[n]) return;
#pragma unroll
    for (int k = 0; k < Q; k++) {
        int xs = x - D2Q9::cx[k], ys = y - D2Q9::cy[k];
        if (g.per_x) { if (xs < 0) xs += g.Nx; else if (xs >= g.Nx) xs -= g.Nx; }
        if (g.per_y) { if (ys < 0) ys += g.Ny; else if (ys >= g.Ny) ys -= g.Ny; }
        if (xs < 0 || xs >= g.Nx || ys < 0 || ys >= g.Ny) continue;
        const size_t ns = (size_t)ys * g.pitch + xs;
        if (solid[ns]) continue;
        D[k * g.plane + n] = S[k * g.plane + ns];
    }
}

// mei_initialization.c:48 -- rho_0 = (sum of rho_in in row-major order) / (Nx*Ny),
// one thread so that the summation order is the reference's.
__global__ void k_seq_mean(const double *__restrict__ R, Geo g, double *__restrict__ out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (int y = 0; y < g.Ny; y++)
        for (int x = 0; x < g.Nx; x++) s += R[(size_t)y * g.pitch + x];
    out[0] = s / (double)(g.Nx * g.Ny);
}

// mei_initialization.c:51-63 -- incompressible equilibrium around rho_0.
__global__ void __launch_bounds__(256) k_mei_feq(double *__restrict__ F0, Fields in, Geo g,
                                                 const double *__restrict__ rho0p)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= g.Nx) return;
    const size_t n = (size_t)y * g.pitch + x;
    const double rho0 = rho0p[0];
    const double ux = in.U[n], uy = in.V[n], r = in.R[n];
    const double usq = ux * ux + uy * uy;
#pragma unroll
    for (int k = 0; k < Q; k++) {
        const double cu = (double)D2Q9::cx[k] * ux + (double)D2Q9::cy[k] * uy;
        F0[k * g.plane + n] =
            D2Q9::w[k] * r + D2Q9::w[k] * rho0 * (3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
    }
}

// mei_initialization.c:72-108 -- save rho, rho = sum f, then collide every node
// (no solid test) towards the incompressible equilibrium with u held fixed.
__global__ void __launch_bounds__(256) k_mei_iter(double *__restrict__ F0, Fields st,
                                                  double *__restrict__ prev, Geo g,
                                                  const double *__restrict__ omega,
                                                  const double *__restrict__ rho0p)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= g.Nx) return;
    const size_t n = (size_t)y * g.pitch + x;
    const double rho0 = rho0p[0];
    prev[n] = st.R[n];
    double f[Q];
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < Q; k++) {
        f[k] = F0[k * g.plane + n];
        s += f[k];
    }
    st.R[n] = s;
    const double ux = st.U[n], uy = st.V[n], om = omega[n];
    const double usq = ux * ux + uy * uy;
#pragma unroll
    for (int k = 0; k < Q; k++) {
        const double cu = (double)D2Q9::cx[k] * ux + (double)D2Q9::cy[k] * uy;
        const double eq =
            D2Q9::w[k] * s + D2Q9::w[k] * rho0 * (3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
        F0[k * g.plane + n] = f[k] + -om * (f[k] - eq);
    }
}

// mei_initialization.c:119-127 -- max |rho - rho_prev| (NaN differences ignored
// as the reference's `dr > max` test ignores them).  Non-negative doubles order
// like their bit patterns, so an integer max is exact.
__global__ void __launch_bounds__(256) k_max_absdiff(const double *__restrict__ a,
                                                     const double *__restrict__ b, Geo g,
                                                     unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long red[256];
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    unsigned long long m = 0;
    if (x < g.Nx) {
        const size_t n = (size_t)y * g.pitch + x;
        const double d = fabs(a[n] - b[n]);
        if (d == d) m = (unsigned long long)__double_as_longlong(d);
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o && red[threadIdx.x + o] > red[threadIdx.x])
            red[threadIdx.x] = red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, red[0]);
}

__global__ void __launch_bounds__(256) k_mask_pitched(int *__restrict__ dst,
                                                      const uint8_t *__restrict__ cls, Geo g)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= g.Nx) return;
    const size_t n = (size_t)y * g.pitch + x;
    dst[n] = (g.geom != 0 && cls[n] == 0) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// cached stage lattice
// ---------------------------------------------------------------------------
namespace {

struct StageCache {
    lbmx_lattice *lat = nullptr;
    int *solid = nullptr;          // pitched int solid flags
    double *omega_raw = nullptr;   // pitched omega of every node (Mei)
    double *scal = nullptr;        // device scalars
    unsigned long long *umax = nullptr;
    int Nx = 0, Ny = 0, per_x = -1, per_y = -1, ibb = -1;
};
StageCache g_stage;

int stage_lattice(int Nx, int Ny, int per_x, int per_y, int ibb, double Fx, double Fy,
                  lbmx_lattice **out)
{
    StageCache &c = g_stage;
    if (!c.lat || c.Nx != Nx || c.Ny != Ny || c.per_x != per_x || c.per_y != per_y ||
        c.ibb != ibb) {
        if (c.lat) lbmx_destroy(c.lat);
        hipFree(c.solid);
        hipFree(c.omega_raw);
        double *scal = c.scal;
        unsigned long long *umax = c.umax;
        c = StageCache();
        c.scal = scal;
        c.umax = umax;
        lbmx_config cfg{};
        cfg.Nx = Nx;
        cfg.Ny = Ny;
        cfg.precision = LBMX_F64;
        cfg.per_x = per_x;
        cfg.per_y = per_y;
        cfg.use_ibb = ibb;
        cfg.flags = LBMX_EXACT_STATE;
        int rc = lbmx_create(&cfg, &c.lat);
        if (rc) return rc;
        const size_t nf = (size_t)Ny * c.lat->pitch;
        HIPCHK(hipMalloc((void **)&c.solid, nf * sizeof(int)));
        HIPCHK(hipMalloc((void **)&c.omega_raw, nf * sizeof(double)));
        if (!c.scal) HIPCHK(hipMalloc((void **)&c.scal, 4 * sizeof(double)));
        if (!c.umax) HIPCHK(hipMalloc((void **)&c.umax, sizeof(unsigned long long)));
        c.Nx = Nx;
        c.Ny = Ny;
        c.per_x = per_x;
        c.per_y = per_y;
        c.ibb = ibb;
    }
    c.lat->cfg.Fx = Fx;
    c.lat->cfg.Fy = Fy;
    c.lat->Fx_lag = Fx;   // a stand-alone stage is its own step: moments use its force
    c.lat->Fy_lag = Fy;
    *out = c.lat;
    return LBMX_OK;
}

int stage_geometry(lbmx_lattice *L, const int *mask, const double *omega, const double *phi)
{
    const size_t nn = (size_t)L->Nx * L->Ny;
    std::vector<double> one;
    if (!omega) {
        one.assign(nn, 1.0);
        omega = one.data();
    }
    int rc = lbmx_set_geometry(L, mask, omega, phi);
    if (rc) return rc;
    const Geo g = lbmx_geo(L);
    dim3 grid((L->Nx + 255) / 256, L->Ny);
    hipLaunchKernelGGL(k_mask_pitched, grid, dim3(256), 0, L->stream, g_stage.solid, L->cls, g);
    HIPCHK(hipGetLastError());
    return LBMX_OK;
}

void stage_flags(lbmx_lattice *L, StepFlags &s)
{
    s = lbmx_flags(L);
    s.lag_arrays = 1;   // stand-alone functions read the caller's rho/u/v
    s.exact = 1;        // and the caller's f as post-collision values
}

}  // namespace

// ---------------------------------------------------------------------------
// stand-alone stages
// ---------------------------------------------------------------------------
extern "C" int lbmx_stage_collision(int Nx, int Ny, double *f, const double *rho,
                                    const double *u, const double *v, const double *omega,
                                    const int *mask, double Fx, double Fy)
{
    lbmx_lattice *L;
    int rc = stage_lattice(Nx, Ny, 0, 0, 0, Fx, Fy, &L);
    if (!rc) rc = stage_geometry(L, mask, omega, nullptr);
    if (!rc) rc = lbmx_put_pops(L, 0, f);
    if (!rc) rc = lbmx_put_field(L, L->R, rho);
    if (!rc) rc = lbmx_put_field(L, L->U, u);
    if (!rc) rc = lbmx_put_field(L, L->V, v);
    if (rc) return rc;
    // the fused kernel in exact-state mode writes the collided populations back
    // in place; its streamed output lands in the scratch buffer and is ignored
    BulkArgs ba;
    ba.g = lbmx_geo(L);
    stage_flags(L, ba.s);
    ba.s.Fx_lag = Fx;
    ba.s.Fy_lag = Fy;
    ba.F = Fields{L->R, L->U, L->V};
    ba.tiles_per_row = (Nx + kBulkThreads - 1) / kBulkThreads;
    ba.ntiles = (long long)ba.tiles_per_row * Ny;
    ba.nxcd = L->nxcd;
    ba.ibb = 1;   // no bounce-back writes into the scratch buffer
    int grid = (int)((ba.ntiles + 7) / 8 * 8);
    if (grid > L->bulk_grid) grid = L->bulk_grid;
    double *S = (double *)L->pop[0], *D = (double *)L->pop[1];
    switch (L->geom) {
    case 0: hipLaunchKernelGGL((k_bulk<double, 0>), dim3(grid), dim3(kBulkThreads), 0, L->stream, S, D, ba); break;
    case 1: hipLaunchKernelGGL((k_bulk<double, 1>), dim3(grid), dim3(kBulkThreads), 0, L->stream, S, D, ba); break;
    default: hipLaunchKernelGGL((k_bulk<double, 2>), dim3(grid), dim3(kBulkThreads), 0, L->stream, S, D, ba); break;
    }
    HIPCHK(hipGetLastError());
    return lbmx_get_pops(L, 0, f);
}

extern "C" int lbmx_stage_streaming(int Nx, int Ny, const double *f, double *f_new,
                                    const int *mask, int per_x, int per_y)
{
    lbmx_lattice *L;
    int rc = stage_lattice(Nx, Ny, per_x, per_y, 0, 0.0, 0.0, &L);
    if (!rc) rc = stage_geometry(L, mask, nullptr, nullptr);
    if (!rc) rc = lbmx_put_pops(L, 0, f);
    if (!rc) rc = lbmx_put_pops(L, 1, f_new);
    if (rc) return rc;
    dim3 grid((Nx + 255) / 256, Ny);
    hipLaunchKernelGGL(k_stream_pull, grid, dim3(256), 0, L->stream, (const double *)L->pop[0],
                       (double *)L->pop[1], lbmx_geo(L), (const int *)g_stage.solid);
    HIPCHK(hipGetLastError());
    return lbmx_get_pops(L, 1, f_new);
}

static int stage_links(lbmx_lattice *L, int mode_write_bb, int read_val)
{
    if (!L->nlinks) return LBMX_OK;
    LinkArgs la;
    la.g = lbmx_geo(L);
    stage_flags(L, la.s);
    la.F = Fields{L->R, L->U, L->V};
    la.links = L->links;
    la.nlinks = L->nlinks;
    la.ibb = L->cfg.use_ibb && !read_val;
    la.write_bb = mode_write_bb;
    la.read_val = read_val;
    la.partials = L->partials;
    hipLaunchKernelGGL(k_links<double>, dim3(L->npartials), dim3(kLinkThreads), 0, L->stream,
                       (const double *)L->pop[0], (double *)L->pop[1], la);
    HIPCHK(hipGetLastError());
    return LBMX_OK;
}

extern "C" int lbmx_stage_bounce_back(int Nx, int Ny, const double *f, double *f_new,
                                      const int *mask, int per_x, int per_y, int use_ibb,
                                      const double *phi)
{
    lbmx_lattice *L;
    int rc = stage_lattice(Nx, Ny, per_x, per_y, use_ibb ? 1 : 0, 0.0, 0.0, &L);
    if (!rc) rc = stage_geometry(L, mask, nullptr, use_ibb ? phi : nullptr);
    if (!rc) rc = lbmx_put_pops(L, 0, f);
    if (!rc) rc = lbmx_put_pops(L, 1, f_new);
    if (!rc) rc = stage_links(L, use_ibb ? 0 : 1, 0);
    if (rc) return rc;
    return lbmx_get_pops(L, 1, f_new);
}

extern "C" int lbmx_stage_force(int Nx, int Ny, const double *f, const double *f_new,
                                const int *mask, int per_x, int per_y, double *Fx, double *Fy)
{
    lbmx_lattice *L;
    int rc = stage_lattice(Nx, Ny, per_x, per_y, 0, 0.0, 0.0, &L);
    if (!rc) rc = stage_geometry(L, mask, nullptr, nullptr);
    if (!rc) rc = lbmx_put_pops(L, 0, f);
    if (!rc) rc = lbmx_put_pops(L, 1, f_new);
    if (!rc) rc = stage_links(L, 0, 1);
    if (rc) return rc;
    if (L->nlinks) {
        BcArgs bc{};
        bc.g = lbmx_geo(L);
        stage_flags(L, bc.s);
        bc.F = Fields{L->R, L->U, L->V};
        bc.nbc = 0;
        bc.partials = L->partials;
        bc.npartials = L->npartials;
        bc.force_out = L->force_dev;
        hipLaunchKernelGGL(k_bc<double>, dim3(1), dim3(kBcThreads), 0, L->stream,
                           (const double *)L->pop[0], (double *)L->pop[1], bc);
        HIPCHK(hipGetLastError());
    }
    double lift = 0.0, drag = 0.0;
    rc = lbmx_forces(L, &lift, &drag);
    if (rc) return rc;
    *Fx = drag;
    *Fy = lift;
    return LBMX_OK;
}

extern "C" int lbmx_stage_moments(int Nx, int Ny, const double *f_new, const int *mask,
                                  double *rho, double *u, double *v, double Fx, double Fy)
{
    lbmx_lattice *L;
    int rc = stage_lattice(Nx, Ny, 0, 0, 0, Fx, Fy, &L);
    if (!rc) rc = stage_geometry(L, mask, nullptr, nullptr);
    if (!rc) rc = lbmx_put_pops(L, L->cur, f_new);
    if (!rc) rc = lbmx_launch_moments(L);
    if (!rc) rc = lbmx_get_field(L, rho, L->Ro);
    if (!rc) rc = lbmx_get_field(L, u, L->Uo);
    if (!rc) rc = lbmx_get_field(L, v, L->Vo);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(L->stream));
    return LBMX_OK;
}

extern "C" int lbmx_stage_bc(int Nx, int Ny, int bc_type, int index, const double *val1,
                             const double *val2, const double *f, double *f_new,
                             const double *rho, const double *u, const double *v,
                             const int *mask)
{
    int rc = lbmx_bc_check(bc_type, index, Nx, Ny);
    if (rc) return rc;
    lbmx_lattice *L;
    rc = stage_lattice(Nx, Ny, 0, 0, 0, 0.0, 0.0, &L);
    if (!rc) rc = stage_geometry(L, mask, nullptr, nullptr);
    if (!rc) rc = lbmx_put_pops(L, 0, f);
    if (!rc) rc = lbmx_put_pops(L, 1, f_new);
    if (!rc) rc = lbmx_put_field(L, L->R, rho);
    if (!rc) rc = lbmx_put_field(L, L->U, u);
    if (!rc) rc = lbmx_put_field(L, L->V, v);
    // a one-entry table; periodic entries are never dropped here because the
    // stage lattice does not wrap in streaming
    if (!rc) rc = lbmx_set_boundaries(L, &bc_type, &index, 1);
    if (!rc) rc = lbmx_set_bc_values(L, 0, val1, val2);
    if (rc) return rc;
    BcArgs bc{};
    bc.g = lbmx_geo(L);
    stage_flags(L, bc.s);
    bc.F = Fields{L->R, L->U, L->V};
    bc.bcs = L->bc_dev;
    bc.nbc = L->nprog;
    bc.npartials = 0;
    bc.force_out = L->force_dev;
    hipLaunchKernelGGL(k_bc<double>, dim3(1), dim3(kBcThreads), 0, L->stream,
                       (const double *)L->pop[0], (double *)L->pop[1], bc);
    HIPCHK(hipGetLastError());
    return lbmx_get_pops(L, 1, f_new);
}

// ---------------------------------------------------------------------------
// Mei consistent initialization (mei_initialization.c:7-142)
// ---------------------------------------------------------------------------
extern "C" int lbmx_mei_initialization(int Nx, int Ny, double *f, double *f_new,
                                       const double *rho_in, const double *U_in,
                                       const double *V_in, double *rho, double *u, double *v,
                                       const double *omega, int per_x, int per_y,
                                       int init_iter_max, double tol_rho, const int *mask)
{
    if (!f || !f_new || !rho_in || !U_in || !V_in || !rho || !u || !v || !omega || !mask)
        return lbmx_fail(LBMX_ERR_ARG, "null argument");
    lbmx_lattice *L;
    int rc = stage_lattice(Nx, Ny, per_x, per_y, 0, 0.0, 0.0, &L);
    if (!rc) rc = stage_geometry(L, mask, omega, nullptr);
    if (!rc) rc = lbmx_put_pops(L, 0, f);
    if (!rc) rc = lbmx_put_pops(L, 1, f_new);
    if (!rc) rc = lbmx_put_field(L, L->R, rho_in);
    if (!rc) rc = lbmx_put_field(L, L->U, U_in);
    if (!rc) rc = lbmx_put_field(L, L->V, V_in);
    if (!rc) rc = lbmx_put_field(L, g_stage.omega_raw, omega);
    if (!rc && !L->Ro) rc = lbmx_launch_moments(L);   // allocates the spare arrays
    if (rc) return rc;
    const Geo g = lbmx_geo(L);
    const Fields st{L->R, L->U, L->V};
    double *prev = L->Ro;
    dim3 grid((Nx + 255) / 256, Ny);
    hipLaunchKernelGGL(k_seq_mean, dim3(1), dim3(64), 0, L->stream, (const double *)L->R, g,
                       g_stage.scal);
    hipLaunchKernelGGL(k_mei_feq, grid, dim3(256), 0, L->stream, (double *)L->pop[0], st, g,
                       (const double *)g_stage.scal);
    HIPCHK(hipGetLastError());

    int a = 0, iters = 0;
    double max_d = 0.0;
    for (int it = 0; it < init_iter_max; it++) {
        iters = it + 1;
        hipLaunchKernelGGL(k_mei_iter, grid, dim3(256), 0, L->stream, (double *)L->pop[a], st,
                           prev, g, (const double *)g_stage.omega_raw,
                           (const double *)g_stage.scal);
        hipLaunchKernelGGL(k_stream_pull, grid, dim3(256), 0, L->stream,
                           (const double *)L->pop[a], (double *)L->pop[a ^ 1], g,
                           (const int *)g_stage.solid);
        a ^= 1;
        HIPCHK(hipMemsetAsync(g_stage.umax, 0, sizeof(unsigned long long), L->stream));
        hipLaunchKernelGGL(k_max_absdiff, grid, dim3(256), 0, L->stream, (const double *)L->R,
                           (const double *)prev, g, g_stage.umax);
        HIPCHK(hipGetLastError());
        unsigned long long bits = 0;
        HIPCHK(hipMemcpyAsync(&bits, g_stage.umax, sizeof bits, hipMemcpyDeviceToHost,
                              L->stream));
        HIPCHK(hipStreamSynchronize(L->stream));
        memcpy(&max_d, &bits, sizeof max_d);
        if (max_d <= tol_rho) break;
    }
    // the reference swaps only its local pointers: the caller's f is buffer 0
    rc = lbmx_get_pops(L, 0, f);
    if (!rc) rc = lbmx_get_pops(L, 1, f_new);
    if (!rc) rc = lbmx_get_field(L, rho, L->R);
    if (!rc) rc = lbmx_get_field(L, u, L->U);
    if (!rc) rc = lbmx_get_field(L, v, L->V);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(L->stream));
    return iters;
}

// Releases the cached stage lattice (tests and long-lived hosts).
extern "C" void lbmx_stage_release(void)
{
    if (g_stage.lat) lbmx_destroy(g_stage.lat);
    hipFree(g_stage.solid);
    hipFree(g_stage.omega_raw);
    hipFree(g_stage.scal);
    hipFree(g_stage.umax);
    g_stage = StageCache();
}
