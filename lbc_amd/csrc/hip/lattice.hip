// lattice.hip -- the resident lattice behind include/lbmx.h.
//
// Host-side work here is set-up only: geometry compilation (node classes, the
// solid-link list with its wall fractions q), boundary-table compilation and
// buffer management.  Every per-step operation runs on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "lbmx.h"
#include "lattice_impl.h"
#include "step_kernels.h"
#include "totals_kernels.h"

using namespace lbmx;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

int lbmx_fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" const char *lbmx_last_error(void) { return g_err.c_str(); }

extern "C" const char *lbmx_device_name(void)
{
    static std::string name;
    if (name.empty()) {
        int dev = 0;
        hipDeviceProp_t p;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess)
            name = std::string(p.name) + " (" + p.gcnArchName + ")";
    }
    return name.c_str();
}

// ---------------------------------------------------------------------------
// copy ceiling: one 16-byte load and one non-temporal 16-byte store per thread
// ---------------------------------------------------------------------------
typedef double dbl2 __attribute__((ext_vector_type(2)));

__global__ void __launch_bounds__(256) k_copy16(const dbl2 *__restrict__ a,
                                                dbl2 *__restrict__ b, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(a[i], &b[i]);
}

extern "C" int lbmx_copy_rate(double gib, double *gbs)
{
    if (!gbs || !(gib > 0.0) || gib > 64.0) return lbmx_fail(LBMX_ERR_ARG, "gib must be in (0, 64]");
    const long long n = (long long)(gib * (double)(1ll << 30)) / 16 / 256 * 256;
    if (n <= 0) return lbmx_fail(LBMX_ERR_ARG, "gib too small");
    dbl2 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = LBMX_OK;
    auto run = [&]() -> int {
        HIPCHK(hipMalloc((void **)&a, n * sizeof(dbl2)));
        HIPCHK(hipMalloc((void **)&b, n * sizeof(dbl2)));
        HIPCHK(hipMemset(a, 0, n * sizeof(dbl2)));
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        const unsigned grid = (unsigned)(n / 256);
        const int reps = 5;
        for (int r = 0; r < 2; r++) hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, 0, a, b, n);
        HIPCHK(hipEventRecord(e0, 0));
        for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, 0, a, b, n);
        HIPCHK(hipEventRecord(e1, 0));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipGetLastError());
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        *gbs = 2.0 * 16.0 * (double)n * reps / ((double)ms * 1e-3) / 1e9;
        return LBMX_OK;
    };
    rc = run();
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return rc;
}

int lbmx_device_info(int *ncu)
{
    static int cached = 0;
    if (!cached) {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        hipDeviceProp_t p;
        HIPCHK(hipGetDeviceProperties(&p, dev));
        cached = p.multiProcessorCount;
    }
    *ncu = cached;
    return LBMX_OK;
}

// ---------------------------------------------------------------------------
// boundary-table helpers
// ---------------------------------------------------------------------------
int lbmx_bc_on_column(int t)
{
    return t == LBMX_BC_CONVECTIVE_INLET || t == LBMX_BC_CONVECTIVE_OUTLET ||
           t == LBMX_BC_PRESSURE_INLET || t == LBMX_BC_PRESSURE_OUTLET ||
           t == LBMX_BC_VELOCITY_INLET || t == LBMX_BC_VELOCITY_OUTLET ||
           t == LBMX_BC_VELOCITY_OUTLET_REG || t == LBMX_BC_VELOCITY_INLET_REG ||
           t == LBMX_BC_PRESSURE_OUTLET_REG || t == LBMX_BC_PRESSURE_INLET_REG;
}

// 0 = no values, 1 = val1 only, 2 = val1 and val2 (which the reference reads)
int lbmx_bc_nvals(int t)
{
    switch (t) {
    case LBMX_BC_VELOCITY_INLET: case LBMX_BC_VELOCITY_OUTLET: case LBMX_BC_VELOCITY_TOP:
    case LBMX_BC_VELOCITY_BOTTOM: case LBMX_BC_VELOCITY_OUTLET_REG:
    case LBMX_BC_VELOCITY_INLET_REG: case LBMX_BC_VELOCITY_TOP_REG:
    case LBMX_BC_VELOCITY_BOTTOM_REG:
        return 2;
    case LBMX_BC_PRESSURE_INLET: case LBMX_BC_PRESSURE_OUTLET:
    case LBMX_BC_PRESSURE_OUTLET_REG: case LBMX_BC_PRESSURE_INLET_REG:
    case LBMX_BC_PRESSURE_TOP_REG: case LBMX_BC_PRESSURE_BOTTOM_REG:
        return 1;
    default:
        return 0;
    }
}

// Range check of index so that every neighbour a BC reads lies in the domain.
int lbmx_bc_check(int t, int index, int Nx, int Ny)
{
    if (t < 1 || t > LBMX_BC_MAX) return lbmx_fail(LBMX_ERR_ARG, "unknown boundary type %d", t);
    if (t == LBMX_BC_PERIODIC_X || t == LBMX_BC_PERIODIC_Y) return LBMX_OK;
    const int L = lbmx_bc_on_column(t) ? Nx : Ny;
    int lo = 0, hi = L - 1;
    switch (t) {
    case LBMX_BC_CONVECTIVE_INLET: case LBMX_BC_PRESSURE_INLET:
    case LBMX_BC_VELOCITY_INLET_REG: case LBMX_BC_PRESSURE_INLET_REG:
    case LBMX_BC_VELOCITY_BOTTOM_REG: case LBMX_BC_PRESSURE_BOTTOM_REG:
    case LBMX_BC_CONVECTIVE_BOTTOM:
        hi = L - 2; break;
    case LBMX_BC_CONVECTIVE_OUTLET: case LBMX_BC_PRESSURE_OUTLET:
    case LBMX_BC_VELOCITY_OUTLET_REG: case LBMX_BC_PRESSURE_OUTLET_REG:
    case LBMX_BC_VELOCITY_TOP_REG: case LBMX_BC_PRESSURE_TOP_REG:
    case LBMX_BC_CONVECTIVE_TOP:
        lo = 1; break;
    default: break;
    }
    if (index < lo || index > hi)
        return lbmx_fail(LBMX_ERR_ARG, "boundary type %d: index %d outside [%d, %d]", t, index,
                         lo, hi);
    return LBMX_OK;
}

// ---------------------------------------------------------------------------
// lattice lifecycle
// ---------------------------------------------------------------------------
static size_t esize_of(int prec) { return prec == LBMX_F64 ? 8 : 4; }

extern "C" int lbmx_create(const lbmx_config *cfg, lbmx_lattice **out)
{
    if (!cfg || !out) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    *out = nullptr;
    if (cfg->Nx < 3 || cfg->Ny < 3)
        return lbmx_fail(LBMX_ERR_ARG, "lattice %dx%d: both sizes must be >= 3", cfg->Nx,
                         cfg->Ny);
    if (cfg->precision != LBMX_F64 && cfg->precision != LBMX_F32_SHIFTED &&
        cfg->precision != LBMX_F32_WIDE)
        return lbmx_fail(LBMX_ERR_ARG, "unknown precision %d", cfg->precision);
    if ((long long)cfg->Nx * cfg->Ny > (1LL << 31) - (1LL << 24))
        return lbmx_fail(LBMX_ERR_ARG, "lattice too large for 32-bit node indexing");

    lbmx_lattice *L = new lbmx_lattice();
    L->cfg = *cfg;
    L->launch_mode = getenv("LBMX_NO_TEMPORAL") ? LBMX_LAUNCH_SINGLE : LBMX_LAUNCH_AUTO;
    if (L->cfg.precision != LBMX_F64 && L->cfg.rho0 == 0.0) L->cfg.rho0 = 1.0;
    if (L->cfg.precision == LBMX_F64) L->cfg.rho0 = 0.0;
    L->Nx = cfg->Nx;
    L->Ny = cfg->Ny;
    L->esize = esize_of(cfg->precision);
    L->pitch = (cfg->Nx + 63) / 64 * 64;
    // planes are separated by a small gap so that the nine plane bases of a
    // large lattice do not share the same high-order address bits
    L->plane = (size_t)L->Ny * L->pitch + 64 * 3;
    int rc = lbmx_device_info(&L->ncu);
    if (rc) { delete L; return rc; }
    L->nxcd = 8;

    hipError_t e = hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete L;
        return lbmx_fail(LBMX_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    const size_t pop_bytes = (size_t)kQ * L->plane * L->esize;
    const size_t fld_bytes = (size_t)L->Ny * L->pitch * sizeof(double);
    if (hipMalloc(&L->pop[0], pop_bytes) != hipSuccess ||
        hipMalloc(&L->pop[1], pop_bytes) != hipSuccess ||
        hipMalloc((void **)&L->R, fld_bytes) != hipSuccess ||
        hipMalloc((void **)&L->U, fld_bytes) != hipSuccess ||
        hipMalloc((void **)&L->V, fld_bytes) != hipSuccess ||
        hipMalloc((void **)&L->force_dev, kPartial * sizeof(double)) != hipSuccess ||
        hipHostMalloc((void **)&L->force_host, kPartial * sizeof(double), hipHostMallocDefault) !=
            hipSuccess) {
        lbmx_destroy(L);
        return lbmx_fail(LBMX_ERR_NOMEM, "device allocation of %zu bytes failed",
                         2 * pop_bytes + 3 * fld_bytes);
    }
    L->bytes = 2 * pop_bytes + 3 * fld_bytes;
    // population gaps and pitch padding are never read as data; clear once so
    // every byte of the buffers is defined
    hipMemsetAsync(L->pop[0], 0, pop_bytes, L->stream);
    hipMemsetAsync(L->pop[1], 0, pop_bytes, L->stream);
    hipMemsetAsync(L->force_dev, 0, kPartial * sizeof(double), L->stream);
    // default geometry: all fluid, omega = 1, no boundaries
    L->geom = 0;
    L->om_uniform = 1.0;
    L->bulk_grid = L->ncu * 8;
    L->bulk_grid = (L->bulk_grid + 7) / 8 * 8;
    hipStreamSynchronize(L->stream);
    *out = L;
    return LBMX_OK;
}

extern "C" void lbmx_destroy(lbmx_lattice *L)
{
    if (!L) return;
    if (L->stream) hipStreamSynchronize(L->stream);
    hipFree(L->pop[0]);
    hipFree(L->pop[1]);
    hipFree(L->R);
    hipFree(L->U);
    hipFree(L->V);
    hipFree(L->Ro);
    hipFree(L->Uo);
    hipFree(L->Vo);
    hipFree(L->cls);
    hipFree(L->lut);
    hipFree(L->om_arr);
    hipFree(L->links);
    hipFree(L->partials);
    lbmx_chan_free(L);
    hipFree(L->force_dev);
    if (L->frec) hipFree(L->frec);
    hipFree(L->bc_dev);
    for (double *p : L->val1) hipFree(p);
    for (double *p : L->val2) hipFree(p);
    hipFree(L->staging);
    hipFree(L->vtk_dev);
    hipFree(L->tot_dev);
    if (L->vtk_host) hipHostFree(L->vtk_host);
    hipFree(L->tdibc.st);
    hipFree(L->tdibc.last);
    hipFree(L->tdibc.active);
    hipFree(L->tdibc.hist);
    hipFree(L->probes.nodes);
    hipFree(L->probes.hist);
    if (L->force_host) hipHostFree(L->force_host);
    if (L->stream) hipStreamDestroy(L->stream);
    delete L;
}

// ---------------------------------------------------------------------------
// geometry compilation
// ---------------------------------------------------------------------------
static bool wrap_or_reject(int &c, int n, int periodic)
{
    if (periodic) {
        if (c < 0) c += n;
        else if (c >= n) c -= n;
        return true;
    }
    return c >= 0 && c < n;
}

extern "C" int lbmx_set_geometry(lbmx_lattice *L, const int *mask, const double *omega,
                                 const double *phi)
{
    if (!L || !mask || !omega) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (L->cfg.use_ibb && !phi) return lbmx_fail(LBMX_ERR_ARG, "use_ibb requires phi");
    if (int prc = lbmx_ensure_push(L)) return prc;
    HIPCHK(hipStreamSynchronize(L->stream));
    lbmx_chan_free(L);
    const int Nx = L->Nx, Ny = L->Ny, pitch = L->pitch;
    const size_t nn = (size_t)Nx * Ny;
    static const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

    // node classes: distinct fluid omegas (bitwise) -> class 1..255, solid -> 0
    size_t nsolid = 0;
    std::map<uint64_t, int> cls_of;
    std::vector<double> lut(256, 0.0);
    bool fits = true;
    uint64_t last_bits = 0;
    bool have_last = false;
    for (size_t n = 0; n < nn; n++) {
        if (mask[n]) { nsolid++; continue; }
        uint64_t bits;
        memcpy(&bits, &omega[n], 8);
        if (have_last && bits == last_bits) continue;
        last_bits = bits;
        have_last = true;
        if (fits && !cls_of.count(bits)) {
            if (cls_of.size() >= 255) fits = false;
            else {
                int c = (int)cls_of.size() + 1;
                cls_of[bits] = c;
                lut[c] = omega[n];
            }
        }
    }
    hipFree(L->cls); L->cls = nullptr;
    hipFree(L->lut); L->lut = nullptr;
    hipFree(L->om_arr); L->om_arr = nullptr;
    L->nsolid = (long long)nsolid;
    L->hsolid.clear();
    if (nsolid) {
        L->hsolid.resize(nn);
        for (size_t n = 0; n < nn; n++) L->hsolid[n] = mask[n] ? 1 : 0;
    }
    L->tail_planned = 0;
    L->nclasses = (int)cls_of.size();
    L->hlut1 = lut[1];
    if (nsolid == 0 && cls_of.size() == 1) {
        L->geom = 0;
        L->om_uniform = lut[1];
    } else {
        L->geom = fits ? 1 : 2;
        std::vector<uint8_t> cls((size_t)Ny * pitch, 0);
        uint64_t cbits = 0;
        uint8_t ccls = 0;
        bool have_c = false;
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                size_t n = (size_t)j * Nx + i;
                uint8_t c = 0;
                if (!mask[n]) {
                    if (fits) {
                        uint64_t bits;
                        memcpy(&bits, &omega[n], 8);
                        if (!have_c || bits != cbits) {
                            cbits = bits;
                            ccls = (uint8_t)cls_of[bits];
                            have_c = true;
                        }
                        c = ccls;
                    } else {
                        c = 1;
                    }
                }
                cls[(size_t)j * pitch + i] = c;
            }
        HIPCHK(hipMalloc((void **)&L->cls, cls.size()));
        HIPCHK(hipMemcpy(L->cls, cls.data(), cls.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMalloc((void **)&L->lut, 256 * sizeof(double)));
        HIPCHK(hipMemcpy(L->lut, lut.data(), 256 * sizeof(double), hipMemcpyHostToDevice));
        if (!fits) {
            HIPCHK(hipMalloc((void **)&L->om_arr, (size_t)Ny * pitch * sizeof(double)));
            HIPCHK(hipMemcpy2D(L->om_arr, pitch * sizeof(double), omega, Nx * sizeof(double),
                               Nx * sizeof(double), Ny, hipMemcpyHostToDevice));
        }
    }

    // fluid -> solid links, row-major, k ascending (compute_force_mem.c:26-58);
    // q and the behind node as interpolated_bounce_back.c:51-89 derives them
    std::vector<Link> links;
    std::vector<long long> target;   // the link's solid node, y * Nx + x
    const double eps = 1e-12;
    const int px = L->cfg.per_x, py = L->cfg.per_y;
    if (nsolid) {
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                if (mask[(size_t)j * Nx + i]) continue;
                for (int k = 1; k < 9; k++) {
                    int in = i + CX[k], jn = j + CY[k];
                    if (!wrap_or_reject(in, Nx, px) || !wrap_or_reject(jn, Ny, py)) continue;
                    if (!mask[(size_t)jn * Nx + in]) continue;
                    Link l;
                    l.n = (long long)j * pitch + i;
                    l.k = k;
                    l.q = 0.5;
                    l.nb = -1;
                    l.bsolid = 0;
                    l.body = 0;
                    l.pad_ = 0;
                    if (L->cfg.use_ibb) {
                        double den = phi[(size_t)j * Nx + i] - phi[(size_t)jn * Nx + in];
                        if (fabs(den) < eps) den = (den >= 0.0) ? eps : -eps;
                        double q = phi[(size_t)j * Nx + i] / den;
                        if (q < eps) q = eps;
                        if (q > 1.0 - eps) q = 1.0 - eps;
                        l.q = q;
                        int ib = i - CX[k], jb = j - CY[k];
                        if (wrap_or_reject(ib, Nx, px) && wrap_or_reject(jb, Ny, py)) {
                            l.nb = (long long)jb * pitch + ib;
                            l.bsolid = mask[(size_t)jb * Nx + ib] ? 1 : 0;
                        }
                    }
                    links.push_back(l);
                    target.push_back((long long)jn * Nx + in);
                }
            }
    }
    hipFree(L->links); L->links = nullptr;
    hipFree(L->partials); L->partials = nullptr;
    L->nlinks = (int)links.size();
    L->npartials = (L->nlinks + kLinkThreads - 1) / kLinkThreads;
    if (L->nlinks) {
        HIPCHK(hipMalloc((void **)&L->links, links.size() * sizeof(Link)));
        HIPCHK(hipMemcpy(L->links, links.data(), links.size() * sizeof(Link),
                         hipMemcpyHostToDevice));
        HIPCHK(hipMalloc((void **)&L->partials, kPartial * sizeof(double) * L->npartials));
    }
    L->hlinks = std::move(links);
    L->link_target = std::move(target);
    HIPCHK(hipMemsetAsync(L->force_dev, 0, kPartial * sizeof(double), L->stream));

    // stale-edge enumeration (k_edges)
    L->edge_rows = py ? 0 : 2 * Nx;
    L->edge_cols = px ? 0 : (py ? 2 * Ny : 2 * (Ny - 2));
    L->has_geometry = 1;
    HIPCHK(hipStreamSynchronize(L->stream));
    return LBMX_OK;
}

extern "C" int lbmx_set_uniform(lbmx_lattice *L, double omega)
{
    if (!L) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (int prc = lbmx_ensure_push(L)) return prc;
    HIPCHK(hipStreamSynchronize(L->stream));
    hipFree(L->cls); L->cls = nullptr;
    hipFree(L->lut); L->lut = nullptr;
    hipFree(L->om_arr); L->om_arr = nullptr;
    hipFree(L->links); L->links = nullptr;
    hipFree(L->partials); L->partials = nullptr;
    lbmx_chan_free(L);
    L->geom = 0;
    L->om_uniform = omega;
    L->nclasses = 1;
    L->hlut1 = omega;
    L->nsolid = 0;
    L->hsolid.clear();
    L->tail_planned = 0;
    L->nlinks = 0;
    L->npartials = 0;
    L->hlinks.clear();
    L->link_target.clear();
    HIPCHK(hipMemsetAsync(L->force_dev, 0, kPartial * sizeof(double), L->stream));
    L->edge_rows = L->cfg.per_y ? 0 : 2 * L->Nx;
    L->edge_cols = L->cfg.per_x ? 0 : (L->cfg.per_y ? 2 * L->Ny : 2 * (L->Ny - 2));
    L->has_geometry = 1;
    HIPCHK(hipStreamSynchronize(L->stream));
    return LBMX_OK;
}

// ---------------------------------------------------------------------------
// boundary table
// ---------------------------------------------------------------------------
static std::vector<int> program_entries(const int *types, int nbc, int per_x, int per_y);

extern "C" int lbmx_set_boundaries(lbmx_lattice *L, const int *types, const int *index, int nbc)
{
    if (!L || nbc < 0 || (nbc > 0 && (!types || !index)))
        return lbmx_fail(LBMX_ERR_ARG, "bad boundary table");
    for (int b = 0; b < nbc; b++) {
        int rc = lbmx_bc_check(types[b], index[b], L->Nx, L->Ny);
        if (rc) return rc;
    }
    if (int prc = lbmx_ensure_push(L)) return prc;
    HIPCHK(hipStreamSynchronize(L->stream));
    for (double *p : L->val1) hipFree(p);
    for (double *p : L->val2) hipFree(p);
    L->val1.assign(nbc, nullptr);
    L->val2.assign(nbc, nullptr);
    L->types.assign(types, types + nbc);
    L->index.assign(index, index + nbc);
    L->vals_set.assign(nbc, 0);
    lbmx_chan_free(L);   // eligibility, columns and work lists depend on the table

    std::vector<BcDesc> prog;
    std::vector<int> edge;
    L->prog_type.clear();
    L->prog_index.clear();
    L->tail_planned = 0;
    for (int b = 0; b < nbc; b++) {
        const int t = types[b];
        const int nv = lbmx_bc_nvals(t);
        const int len = lbmx_bc_on_column(t) ? L->Ny : L->Nx;
        if (nv >= 1) HIPCHK(hipMalloc((void **)&L->val1[b], len * sizeof(double)));
        if (nv >= 2) HIPCHK(hipMalloc((void **)&L->val2[b], len * sizeof(double)));
    }
    for (int b : program_entries(types, nbc, L->cfg.per_x, L->cfg.per_y)) {
        const int t = types[b];
        BcDesc d;
        d.type = t;
        d.index = index[b];
        d.v1 = L->val1[b];
        d.v2 = L->val2[b];
        prog.push_back(d);
        L->prog_type.push_back(t);
        L->prog_index.push_back(index[b]);
        edge.push_back((lbmx_bc_on_column(t) || t == LBMX_BC_PERIODIC_X) ? L->Ny : L->Nx);
    }
    L->prog_edge = edge;
    L->prog_desc = prog;
    hipFree(L->bc_dev);
    L->bc_dev = nullptr;
    L->nprog = (int)prog.size();
    if (L->nprog) {
        HIPCHK(hipMalloc((void **)&L->bc_dev, prog.size() * sizeof(BcDesc)));
        HIPCHK(hipMemcpy(L->bc_dev, prog.data(), prog.size() * sizeof(BcDesc),
                         hipMemcpyHostToDevice));
    }
    return LBMX_OK;
}

extern "C" int lbmx_set_bc_values(lbmx_lattice *L, int b, const double *v1, const double *v2)
{
    if (!L || b < 0 || b >= (int)L->types.size())
        return lbmx_fail(LBMX_ERR_ARG, "boundary %d out of range", b);
    const int t = L->types[b];
    const int nv = lbmx_bc_nvals(t);
    const int len = lbmx_bc_on_column(t) ? L->Ny : L->Nx;
    if (nv >= 1) {
        if (!v1) return lbmx_fail(LBMX_ERR_ARG, "boundary %d (type %d) needs val1", b, t);
        HIPCHK(hipMemcpyAsync(L->val1[b], v1, len * sizeof(double), hipMemcpyHostToDevice,
                              L->stream));
    }
    if (nv >= 2) {
        if (!v2) return lbmx_fail(LBMX_ERR_ARG, "boundary %d (type %d) needs val2", b, t);
        HIPCHK(hipMemcpyAsync(L->val2[b], v2, len * sizeof(double), hipMemcpyHostToDevice,
                              L->stream));
    }
    L->vals_set[b] = 1;
    return LBMX_OK;
}

// ---------------------------------------------------------------------------
// state transfer
// ---------------------------------------------------------------------------
Geo lbmx_geo(const lbmx_lattice *L)
{
    Geo g;
    g.Nx = L->Nx;
    g.Ny = L->Ny;
    g.pitch = L->pitch;
    g.plane = L->plane;
    g.geom = L->geom;
    g.cls = L->cls;
    g.lut = L->lut;
    g.om_arr = L->om_arr;
    g.om = L->om_uniform;
    g.per_x = L->cfg.per_x;
    g.per_y = L->cfg.per_y;
    g.rho0 = L->cfg.rho0;
    g.wide = L->cfg.precision == LBMX_F32_WIDE;
    return g;
}

static int ensure_staging(lbmx_lattice *L, size_t bytes)
{
    if (L->staging_bytes >= bytes) return LBMX_OK;
    hipFree(L->staging);
    L->staging = nullptr;
    L->staging_bytes = 0;
    HIPCHK(hipMalloc((void **)&L->staging, bytes));
    L->staging_bytes = bytes;
    return LBMX_OK;
}

// rows of AoS populations per transfer chunk (<= 64 MiB of staging)
static int chunk_rows(const lbmx_lattice *L)
{
    size_t row = (size_t)L->Nx * kQ * sizeof(double);
    size_t r = ((size_t)64 << 20) / row;
    if (r < 1) r = 1;
    if (r > (size_t)L->Ny) r = L->Ny;
    return (int)r;
}

int lbmx_put_pops(lbmx_lattice *L, int which, const double *aos)
{
    const int rows = chunk_rows(L);
    int rc = ensure_staging(L, (size_t)rows * L->Nx * kQ * sizeof(double));
    if (rc) return rc;
    Geo g = lbmx_geo(L);
    for (int y0 = 0; y0 < L->Ny; y0 += rows) {
        const int r = (y0 + rows <= L->Ny) ? rows : L->Ny - y0;
        const size_t cnt = (size_t)r * L->Nx;
        HIPCHK(hipMemcpyAsync(L->staging, aos + (size_t)y0 * L->Nx * kQ, cnt * kQ * 8,
                              hipMemcpyHostToDevice, L->stream));
        const unsigned blocks = (unsigned)((cnt + 255) / 256);
        if (L->esize == 8)
            hipLaunchKernelGGL(k_aos_to_planes<double>, dim3(blocks), dim3(256), 0, L->stream,
                               (double *)L->pop[which], L->staging, g, y0, r);
        else
            hipLaunchKernelGGL(k_aos_to_planes<float>, dim3(blocks), dim3(256), 0, L->stream,
                               (float *)L->pop[which], L->staging, g, y0, r);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(L->stream));
    }
    return LBMX_OK;
}

int lbmx_get_pops(lbmx_lattice *L, int which, double *aos)
{
    const int rows = chunk_rows(L);
    int rc = ensure_staging(L, (size_t)rows * L->Nx * kQ * sizeof(double));
    if (rc) return rc;
    Geo g = lbmx_geo(L);
    for (int y0 = 0; y0 < L->Ny; y0 += rows) {
        const int r = (y0 + rows <= L->Ny) ? rows : L->Ny - y0;
        const size_t cnt = (size_t)r * L->Nx;
        const unsigned blocks = (unsigned)((cnt + 255) / 256);
        if (L->esize == 8)
            hipLaunchKernelGGL(k_planes_to_aos<double>, dim3(blocks), dim3(256), 0, L->stream,
                               L->staging, (const double *)L->pop[which], g, y0, r);
        else
            hipLaunchKernelGGL(k_planes_to_aos<float>, dim3(blocks), dim3(256), 0, L->stream,
                               L->staging, (const float *)L->pop[which], g, y0, r);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(aos + (size_t)y0 * L->Nx * kQ, L->staging, cnt * kQ * 8,
                              hipMemcpyDeviceToHost, L->stream));
        HIPCHK(hipStreamSynchronize(L->stream));
    }
    return LBMX_OK;
}

// Dense rows [y0, y0 + r) of a pitched field <-> a contiguous [r][Nx] buffer.
__global__ void __launch_bounds__(256) k_field_rows(double *__restrict__ dense,
                                                    double *__restrict__ field, int pitch, int Nx,
                                                    int y0, long long cnt, int to_dense)
{
    const long long n = (long long)blockIdx.x * 256 + threadIdx.x;
    if (n >= cnt) return;
    const int y = (int)(n / Nx), x = (int)(n - (long long)y * Nx);
    double *f = field + (size_t)(y0 + y) * pitch + x;
    if (to_dense)
        dense[n] = *f;
    else
        *f = dense[n];
}

// Rows of a field per transfer chunk (<= 64 MiB of staging).
static int field_chunk_rows(const lbmx_lattice *L)
{
    size_t r = ((size_t)64 << 20) / ((size_t)L->Nx * sizeof(double));
    if (r < 1) r = 1;
    if (r > (size_t)L->Ny) r = L->Ny;
    return (int)r;
}

// Fields cross PCIe as contiguous copies (a pitched 2D copy to or from pageable
// memory moves row by row and ran at ~2 GB/s): directly when the pitch equals Nx,
// else through dense rows in the staging buffer.
int lbmx_put_field(lbmx_lattice *L, double *dev, const double *host)
{
    const size_t row = (size_t)L->Nx * sizeof(double);
    if (L->pitch == L->Nx) {
        HIPCHK(hipMemcpyAsync(dev, host, row * L->Ny, hipMemcpyHostToDevice, L->stream));
        return LBMX_OK;
    }
    const int rows = field_chunk_rows(L);
    if (int rc = ensure_staging(L, (size_t)rows * row)) return rc;
    for (int y0 = 0; y0 < L->Ny; y0 += rows) {
        const int r = std::min(rows, L->Ny - y0);
        const long long cnt = (long long)r * L->Nx;
        HIPCHK(hipMemcpyAsync(L->staging, host + (size_t)y0 * L->Nx, cnt * sizeof(double),
                              hipMemcpyHostToDevice, L->stream));
        hipLaunchKernelGGL(k_field_rows, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0,
                           L->stream, L->staging, dev, L->pitch, L->Nx, y0, cnt, 0);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(L->stream));   // the staging rows are reused
    }
    return LBMX_OK;
}

int lbmx_get_field(lbmx_lattice *L, double *host, const double *dev)
{
    const size_t row = (size_t)L->Nx * sizeof(double);
    if (L->pitch == L->Nx) {
        HIPCHK(hipMemcpyAsync(host, dev, row * L->Ny, hipMemcpyDeviceToHost, L->stream));
        return LBMX_OK;
    }
    const int rows = field_chunk_rows(L);
    if (int rc = ensure_staging(L, (size_t)rows * row)) return rc;
    for (int y0 = 0; y0 < L->Ny; y0 += rows) {
        const int r = std::min(rows, L->Ny - y0);
        const long long cnt = (long long)r * L->Nx;
        hipLaunchKernelGGL(k_field_rows, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0,
                           L->stream, L->staging, const_cast<double *>(dev), L->pitch, L->Nx, y0,
                           cnt, 1);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(host + (size_t)y0 * L->Nx, L->staging, cnt * sizeof(double),
                              hipMemcpyDeviceToHost, L->stream));
        HIPCHK(hipStreamSynchronize(L->stream));   // the staging rows are reused
    }
    return LBMX_OK;
}

extern "C" int lbmx_upload(lbmx_lattice *L, const double *f, const double *f_new,
                           const double *rho, const double *u, const double *v)
{
    if (!L || !f || !rho || !u || !v) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (!L->has_geometry) return lbmx_fail(LBMX_ERR_STATE, "lbmx_set_geometry first");
    int rc = lbmx_put_pops(L, L->cur, f);
    if (!rc) rc = lbmx_put_pops(L, L->cur ^ 1, f_new ? f_new : f);
    if (!rc) rc = lbmx_put_field(L, L->R, rho);
    if (!rc) rc = lbmx_put_field(L, L->U, u);
    if (!rc) rc = lbmx_put_field(L, L->V, v);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(L->stream));
    lbmx_reset_run(L);
    return LBMX_OK;
}

void lbmx_reset_run(lbmx_lattice *L)
{
    L->lag_arrays = 1;
    L->has_state = 1;
    L->steps = 0;
    L->Fx_lag = L->cfg.Fx;
    L->Fy_lag = L->cfg.Fy;
    L->other_is_ref = 1;
    L->pull = 0;
}

extern "C" int lbmx_set_force(lbmx_lattice *L, double Fx, double Fy)
{
    if (!L) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    // (in gather form the stored field is the last step's post-collision one, so
    // the next collision has not run yet and takes the new force)
    L->cfg.Fx = Fx;
    L->cfg.Fy = Fy;
    return LBMX_OK;
}

extern "C" int lbmx_init_taylor_green(lbmx_lattice *L, double u0, double kx, double ky)
{
    if (!L) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (!L->has_geometry) return lbmx_fail(LBMX_ERR_STATE, "lbmx_set_geometry first");
    if (ky == 0.0) return lbmx_fail(LBMX_ERR_ARG, "ky must be non-zero");
    const Geo g = lbmx_geo(L);
    const Fields F{L->R, L->U, L->V};
    dim3 grid((L->Nx + 255) / 256, L->Ny);
    if (L->esize == 8)
        hipLaunchKernelGGL(k_init_tgv<double>, grid, dim3(256), 0, L->stream,
                           (double *)L->pop[0], (double *)L->pop[1], F, g, u0, kx, ky);
    else
        hipLaunchKernelGGL(k_init_tgv<float>, grid, dim3(256), 0, L->stream,
                           (float *)L->pop[0], (float *)L->pop[1], F, g, u0, kx, ky);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(L->stream));
    lbmx_reset_run(L);
    return LBMX_OK;
}

// Completes the post-collision field in the non-current buffer (fast mode).
int lbmx_materialize_other(lbmx_lattice *L)
{
    if (L->other_is_ref) return LBMX_OK;
    const Geo g = lbmx_geo(L);
    const Fields F{L->R, L->U, L->V};
    dim3 grid((L->Nx + 255) / 256, L->Ny);
    if (L->esize == 8)
        hipLaunchKernelGGL(k_materialize<double>, grid, dim3(256), 0, L->stream,
                           (double *)L->pop[L->cur ^ 1], g, L->last, F);
    else
        hipLaunchKernelGGL(k_materialize<float>, grid, dim3(256), 0, L->stream,
                           (float *)L->pop[L->cur ^ 1], g, L->last, F);
    HIPCHK(hipGetLastError());
    L->other_is_ref = 1;
    return LBMX_OK;
}

extern "C" int lbmx_download(lbmx_lattice *L, double *f, double *f_new)
{
    if (!L) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (!L->has_state) return lbmx_fail(LBMX_ERR_STATE, "no state uploaded");
    int rc = lbmx_ensure_push(L);
    if (!rc && f) rc = lbmx_get_pops(L, L->cur, f);
    if (!rc && f_new) rc = lbmx_materialize_other(L);
    if (!rc && f_new) rc = lbmx_get_pops(L, L->cur ^ 1, f_new);
    return rc;
}

// ---------------------------------------------------------------------------
// time stepping
// ---------------------------------------------------------------------------
StepFlags lbmx_flags(const lbmx_lattice *L)
{
    StepFlags s;
    s.lag_arrays = L->lag_arrays;
    s.exact = (L->cfg.flags & LBMX_EXACT_STATE) ? 1 : 0;
    s.Fx = L->cfg.Fx;
    s.Fy = L->cfg.Fy;
    s.force = (L->cfg.Fx != 0.0 || L->cfg.Fy != 0.0);
    s.Fx_lag = L->Fx_lag;
    s.Fy_lag = L->Fy_lag;
    return s;
}

template <typename T>
static int launch_tdibc(lbmx_lattice *L, const T *S, const Geo &g, const StepFlags &s,
                        const Fields &F);
template <typename T>
static int launch_probes(lbmx_lattice *L, const T *S, const Geo &g, const StepFlags &s,
                         const Fields &F);

// ---------------------------------------------------------------------------
// planning the work after streaming (k_tail)
// ---------------------------------------------------------------------------
// Nodes (y * Nx + x) a program entry writes in D, reads in D and reads in S or in
// the lagged-moment arrays, visited as the bc_* kernels visit them.
namespace {
struct NodeSets {
    std::vector<long long> wd, rd, rs;
};

struct PlanGrid {
    int Nx, Ny, per_x, per_y;
    const uint8_t *solid;   // Ny x Nx, or null: no solid node
};

bool solid_h(const PlanGrid *L, int x, int y)
{
    return L->solid && L->solid[(size_t)y * L->Nx + x];
}

NodeSets entry_nodes(const PlanGrid *L, int t, int idx)
{
    static const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    const int Nx = L->Nx, Ny = L->Ny;
    auto id = [Nx](int x, int y) { return (long long)y * Nx + x; };
    NodeSets s;
    switch (t) {
    case 3: case 4: case 7: case 8: {   // Zou-He columns: bc_zou_he_column
        const bool inlet = (t == 7 || t == 3), pressure = (t == 3 || t == 4);
        for (int j = 0; j < Ny; j++) {
            if (solid_h(L, idx, j)) continue;
            s.wd.push_back(id(idx, j));
            s.rd.push_back(id(idx, j));
            if (pressure) s.rs.push_back(id(idx + (inlet ? 1 : -1), j));
        }
        break;
    }
    case 9: case 10:                    // Zou-He rows: bc_zou_he_row
        for (int i = 0; i < Nx; i++) {
            if (solid_h(L, i, idx)) continue;
            s.wd.push_back(id(i, idx));
            s.rd.push_back(id(i, idx));
        }
        break;
    case 1: case 2: case 19: case 20: { // bc_convective
        const bool column = (t == 1 || t == 2);
        const int from = idx + ((t == 1 || t == 20) ? 1 : -1);
        const int n = column ? Ny : Nx;
        for (int e = 0; e < n; e++) {
            const int xb = column ? idx : e, yb = column ? e : idx;
            if (solid_h(L, xb, yb)) continue;
            s.wd.push_back(id(xb, yb));
            s.rd.push_back(column ? id(from, e) : id(e, from));
        }
        break;
    }
    case 5: case 6: {                   // bc_periodic
        const bool along_x = (t == 5);
        const int n = along_x ? Ny : Nx, len = along_x ? Nx : Ny;
        for (int e = 0; e < n; e++)
            for (int k = 1; k < 9; k++) {
                const int c = along_x ? CX[k] : CY[k];
                if (c == 0) continue;
                const int dst = c > 0 ? 0 : len - 1, src = c > 0 ? len - c : -c - 1;
                int i, j, is, js;
                if (along_x) {
                    i = dst; j = e; is = src; js = e - CY[k];
                    if (js < 0 || js >= Ny) continue;
                } else {
                    i = e; j = dst; is = e - CX[k]; js = src;
                    if (is < 0 || is >= Nx) continue;
                }
                if (solid_h(L, i, j) || solid_h(L, is, js)) continue;
                s.wd.push_back(id(i, j));
                s.rs.push_back(id(is, js));
            }
        break;
    }
    default: {                          // bc_regularized
        const bool column = (t == 11 || t == 12 || t == 15 || t == 16);
        const int step_in = (t == 12 || t == 14 || t == 16 || t == 18) ? 1 : -1;
        const int n = column ? Ny : Nx;
        for (int e = 0; e < n; e++) {
            const int xb = column ? idx : e, yb = column ? e : idx;
            if (solid_h(L, xb, yb)) continue;
            const long long ni = column ? id(idx + step_in, e) : id(e, idx + step_in);
            s.wd.push_back(id(xb, yb));
            s.rd.push_back(ni);
            s.rs.push_back(ni);
        }
        break;
    }
    }
    for (auto *v : {&s.wd, &s.rd, &s.rs}) {
        std::sort(v->begin(), v->end());
        v->erase(std::unique(v->begin(), v->end()), v->end());
    }
    return s;
}

bool meet(const std::vector<long long> &a, const std::vector<long long> &b)
{
    size_t i = 0, j = 0;
    while (i < a.size() && j < b.size()) {
        if (a[i] == b[j]) return true;
        if (a[i] < b[j]) i++;
        else j++;
    }
    return false;
}

// The stale-edge job's nodes: fluid nodes of the non-periodic domain edges, whose
// S slots and lagged moments edge_node rewrites.
std::vector<long long> edge_nodes(const PlanGrid *L)
{
    const int Nx = L->Nx, Ny = L->Ny;
    std::vector<long long> v;
    auto add = [&](int x, int y) {
        if (!solid_h(L, x, y)) v.push_back((long long)y * Nx + x);
    };
    if (!L->per_y)
        for (int i = 0; i < Nx; i++) { add(i, 0); add(i, Ny - 1); }
    if (!L->per_x)
        for (int j = 0; j < Ny; j++) { add(0, j); add(Nx - 1, j); }
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    return v;
}

// Consecutive program entries share a launch while no entry writes a node that an
// earlier one of the launch reads or writes, or reads a node that one writes; the
// edge job (edges: the domain has a non-periodic axis) joins the last launch when
// no entry there reads S on an edge node.
std::vector<lbmx_lattice::TailLaunch> plan_tail(const PlanGrid *G, const int *type,
                                                const int *index, int n, bool edges)
{
    std::vector<lbmx_lattice::TailLaunch> plan;
    std::vector<NodeSets> cur;
    lbmx_lattice::TailLaunch launch;
    for (int b = 0; b < n; b++) {
        NodeSets s = entry_nodes(G, type[b], index[b]);
        bool clash = (int)launch.entry.size() >= kTailJobs - 1;
        for (const NodeSets &o : cur) {
            if (clash) break;
            clash = meet(o.wd, s.wd) || meet(o.wd, s.rd) || meet(o.rd, s.wd);
        }
        if (clash) {
            plan.push_back(launch);
            launch = lbmx_lattice::TailLaunch{};
            cur.clear();
        }
        launch.entry.push_back(b);
        cur.push_back(std::move(s));
    }
    if (edges) {
        const std::vector<long long> en = edge_nodes(G);
        bool own_launch = launch.entry.empty();
        for (const NodeSets &o : cur)
            if (!own_launch) own_launch = meet(o.rs, en);
        if (own_launch) {
            if (!launch.entry.empty()) plan.push_back(launch);
            launch = lbmx_lattice::TailLaunch{};
        }
        launch.edges = 1;
    }
    if (!launch.entry.empty() || launch.edges) plan.push_back(launch);
    return plan;
}
}   // namespace

// Leading periodic entries whose axis already wraps in streaming re-write exactly
// what streaming wrote (periodic_x.c vs streaming.c:24-34): they are left out of
// the program.  Returns the table indices of the program's entries.
static std::vector<int> program_entries(const int *types, int nbc, int per_x, int per_y)
{
    std::vector<int> keep;
    for (int b = 0; b < nbc; b++) {
        const bool redundant = (types[b] == LBMX_BC_PERIODIC_X && per_x) ||
                               (types[b] == LBMX_BC_PERIODIC_Y && per_y);
        if (redundant && keep.empty()) continue;
        keep.push_back(b);
    }
    return keep;
}

static void lbmx_plan_tail(lbmx_lattice *L)
{
    const PlanGrid G{L->Nx, L->Ny, L->cfg.per_x, L->cfg.per_y,
                     L->hsolid.empty() ? nullptr : L->hsolid.data()};
    L->tail = plan_tail(&G, L->prog_type.data(), L->prog_index.data(), L->nprog,
                        L->edge_rows + L->edge_cols > 0);
    L->tail_planned = 1;
}

extern "C" int lbmx_tail_plan(int Nx, int Ny, int per_x, int per_y, const int *mask,
                              const int *types, const int *index, int nbc, int *launch_of,
                              int *edge_launch)
{
    if (Nx < 3 || Ny < 3 || nbc < 0 || (nbc > 0 && (!types || !index || !launch_of)) ||
        !edge_launch)
        return lbmx_fail(LBMX_ERR_ARG, "invalid argument");
    for (int b = 0; b < nbc; b++) {
        int rc = lbmx_bc_check(types[b], index[b], Nx, Ny);
        if (rc) return rc;
    }
    std::vector<uint8_t> solid;
    if (mask) {
        solid.resize((size_t)Nx * Ny);
        for (size_t n = 0; n < solid.size(); n++) solid[n] = mask[n] ? 1 : 0;
    }
    const PlanGrid G{Nx, Ny, per_x ? 1 : 0, per_y ? 1 : 0, mask ? solid.data() : nullptr};
    const std::vector<int> keep = program_entries(types, nbc, per_x, per_y);
    std::vector<int> pt, pi;
    for (int b : keep) {
        pt.push_back(types[b]);
        pi.push_back(index[b]);
    }
    const auto plan = plan_tail(&G, pt.data(), pi.data(), (int)keep.size(), !(per_x && per_y));
    for (int b = 0; b < nbc; b++) launch_of[b] = -1;
    *edge_launch = -1;
    for (size_t l = 0; l < plan.size(); l++) {
        for (int e : plan[l].entry) launch_of[keep[e]] = (int)l;
        if (plan[l].edges) *edge_launch = (int)l;
    }
    return (int)plan.size();
}

template <typename T, bool W>
static void launch_aligned(lbmx_lattice *L, T *S, T *D, const BulkArgs &ba, const LinkArgs &la,
                           int nl, bool fused, unsigned blocks)
{
    const dim3 grid(blocks), block(kAlignThreads);
    switch (L->geom) {
    case 0:
        hipLaunchKernelGGL((k_bulk_al<T, 0, W>), grid, block, 0, L->stream, S, D, ba);
        break;
    case 1:
        if (fused)
            hipLaunchKernelGGL((k_bulk_links_al<T, 1, W>), grid, block, 0, L->stream, S, D, ba,
                               la, nl);
        else
            hipLaunchKernelGGL((k_bulk_al<T, 1, W>), grid, block, 0, L->stream, S, D, ba);
        break;
    default:
        if (fused)
            hipLaunchKernelGGL((k_bulk_links_al<T, 2, W>), grid, block, 0, L->stream, S, D, ba,
                               la, nl);
        else
            hipLaunchKernelGGL((k_bulk_al<T, 2, W>), grid, block, 0, L->stream, S, D, ba);
        break;
    }
}

template <typename T>
static int launch_step(lbmx_lattice *L)
{
    T *S = (T *)L->pop[L->cur];
    T *D = (T *)L->pop[L->cur ^ 1];
    const Geo g = lbmx_geo(L);
    const StepFlags s = lbmx_flags(L);
    const Fields F{L->R, L->U, L->V};

    if (L->probes.n) {
        int rc = launch_probes<T>(L, S, g, s, F);
        if (rc) return rc;
    }
    if (L->tdibc.on) {
        int rc = launch_tdibc<T>(L, S, g, s, F);
        if (rc) return rc;
    }

    BulkArgs ba;
    ba.g = g;
    ba.s = s;
    ba.F = F;
    ba.tiles_per_row = (L->Nx + kBulkThreads - 1) / kBulkThreads;
    ba.ntiles = (long long)ba.tiles_per_row * L->Ny;
    ba.nxcd = L->nxcd;
    ba.ibb = L->cfg.use_ibb;
    int grid = L->bulk_grid;
    if ((long long)grid > ba.ntiles) grid = (int)((ba.ntiles + 7) / 8 * 8);
    if (grid < 8) grid = 8;

    LinkArgs la;
    la.g = g;
    la.s = s;
    la.F = F;
    la.links = L->links;
    la.nlinks = L->nlinks;
    la.ibb = L->cfg.use_ibb;
    la.write_bb = 0;
    la.read_val = 0;
    la.partials = L->partials;
    // links in the bulk launch unless the bulk rewrites S (k_bulk_links)
    const bool fused = L->nlinks > 0 && !s.exact;
    const int link_grid = (L->npartials + L->nxcd - 1) / L->nxcd * L->nxcd;
    const int gridA = grid + (fused ? link_grid : 0);
    static int scatter = -1;   // LBMX_PUSH_SCATTER=1: the shifted-store k_bulk in fast mode too
    if (scatter < 0) scatter = getenv("LBMX_PUSH_SCATTER") ? 1 : 0;
    if (!s.exact && !scatter) {
        // destination-aligned push, one row tile per workgroup (k_bulk_al)
        const long long tiles = ba.ntiles;
        const int nl = fused ? L->npartials : 0;
        if (tiles + nl > 0x7fffffffLL) return lbmx_fail(LBMX_ERR_ARG, "lattice too large");
        if (sizeof(T) == 4 && g.wide)
            launch_aligned<T, true>(L, S, D, ba, la, nl, fused, (unsigned)(tiles + nl));
        else
            launch_aligned<T, false>(L, S, D, ba, la, nl, fused, (unsigned)(tiles + nl));
    } else switch (L->geom) {
    case 0:
        hipLaunchKernelGGL((k_bulk<T, 0>), dim3(grid), dim3(kBulkThreads), 0, L->stream, S, D, ba);
        break;
    case 1:
        if (fused)
            hipLaunchKernelGGL((k_bulk_links<T, 1>), dim3(gridA), dim3(kBulkThreads), 0,
                               L->stream, S, D, ba, la, L->npartials, link_grid);
        else
            hipLaunchKernelGGL((k_bulk<T, 1>), dim3(grid), dim3(kBulkThreads), 0, L->stream, S,
                               D, ba);
        break;
    default:
        if (fused)
            hipLaunchKernelGGL((k_bulk_links<T, 2>), dim3(gridA), dim3(kBulkThreads), 0,
                               L->stream, S, D, ba, la, L->npartials, link_grid);
        else
            hipLaunchKernelGGL((k_bulk<T, 2>), dim3(grid), dim3(kBulkThreads), 0, L->stream, S,
                               D, ba);
        break;
    }
    HIPCHK(hipGetLastError());
    if (L->nlinks && !fused) {
        hipLaunchKernelGGL(k_links<T>, dim3(L->npartials), dim3(kLinkThreads), 0, L->stream,
                           (const T *)S, D, la);
        HIPCHK(hipGetLastError());
    }

    // boundary table, force outputs and stale edges: k_tail launches as planned
    if (!L->tail_planned) lbmx_plan_tail(L);
    TailArgs ta{};
    ta.bc.g = g;
    ta.bc.s = s;
    ta.bc.F = F;
    ta.bc.bcs = L->bc_dev;
    ta.bc.nbc = L->nprog;
    ta.bc.partials = L->partials;
    ta.bc.npartials = L->nlinks ? L->npartials : 0;
    ta.bc.force_out = L->force_dev;
    ta.bc.record = L->frec;
    ta.bc.row = L->frec_row;
    ta.bc.capacity = L->frec_cap;
    ta.bc.dt = L->frec_dt;
    ta.bc.cf = L->frec_cf;
    ta.nrow = L->edge_rows;
    ta.ncol = L->edge_cols;
    bool force_due = L->nlinks > 0 || L->frec;
    if (L->frec) L->frec_row++;
    for (const auto &tl : L->tail) {
        ta.njobs = 0;
        ta.block0[0] = 0;
        for (int b : tl.entry) {
            ta.entry[ta.njobs] = b;
            ta.desc[ta.njobs] = L->prog_desc[b];
            ta.block0[ta.njobs + 1] =
                ta.block0[ta.njobs] + (L->prog_edge[b] + kTailThreads - 1) / kTailThreads;
            ta.njobs++;
        }
        if (tl.edges && !s.exact) {
            ta.entry[ta.njobs] = -1;
            ta.block0[ta.njobs + 1] =
                ta.block0[ta.njobs] + (ta.nrow + ta.ncol + kTailThreads - 1) / kTailThreads;
            ta.njobs++;
        }
        if (ta.njobs == 0) continue;
        ta.with_force = force_due ? 1 : 0;
        force_due = false;
        hipLaunchKernelGGL(k_tail<T>, dim3(ta.block0[ta.njobs]), dim3(kTailThreads), 0,
                           L->stream, S, D, ta);
        HIPCHK(hipGetLastError());
    }
    if (force_due) {   // links or a record but nothing else after streaming
        ta.njobs = 0;
        ta.block0[0] = 0;
        ta.with_force = 1;
        hipLaunchKernelGGL(k_tail<T>, dim3(1), dim3(kTailThreads), 0, L->stream, S, D, ta);
        HIPCHK(hipGetLastError());
    }
    L->cur ^= 1;
    L->last = s;
    L->other_is_ref = s.exact;
    L->lag_arrays = 0;
    L->Fx_lag = L->cfg.Fx;
    L->Fy_lag = L->cfg.Fy;
    L->steps++;
    return LBMX_OK;
}

// ---------------------------------------------------------------------------
// gather form for periodic all-fluid lattices (pull_kernels.h)
// ---------------------------------------------------------------------------
static bool pull_eligible(const lbmx_lattice *L)
{
    static int disabled = -1;
    if (disabled < 0) disabled = getenv("LBMX_NO_GATHER") ? 1 : 0;
    return !disabled && L->geom == 0 && L->cfg.per_x && L->cfg.per_y && L->nprog == 0 &&
           !L->tdibc.on && L->probes.n == 0 && !L->frec && !(L->cfg.flags & LBMX_EXACT_STATE);
}

static PullArgs pull_args(const lbmx_lattice *L)
{
    PullArgs a;
    a.g = lbmx_geo(L);
    a.s = lbmx_flags(L);
    a.F = Fields{L->R, L->U, L->V};
    a.tiles_per_row = (L->Nx + kPullThreads - 1) / kPullThreads;
    return a;
}

// Entering the gather form is the collision of the next step, in place: on a
// periodic all-fluid lattice the step's streaming is the permutation the next
// gather applies, so the buffer then holds that step's post-collision field (the
// reference's f_new after the driver's swap) and the step counts as done.
template <typename T>
static int enter_pull(lbmx_lattice *L)
{
    const PullArgs a = pull_args(L);
    dim3 grid((L->Nx + 255) / 256, L->Ny);
    hipLaunchKernelGGL(k_collide_inplace<T>, grid, dim3(256), 0, L->stream, (T *)L->pop[L->cur],
                       a);
    HIPCHK(hipGetLastError());
    L->pull = 1;
    L->pstar = L->cur;
    L->lag_arrays = 0;
    L->last = a.s;
    L->Fx_lag = L->cfg.Fx;
    L->Fy_lag = L->cfg.Fy;
    L->steps++;
    return LBMX_OK;
}

template <typename T>
static int pull_step(lbmx_lattice *L)
{
    const PullArgs a = pull_args(L);
    const unsigned blocks = (unsigned)a.tiles_per_row * (unsigned)L->Ny;
    const T *S = (const T *)L->pop[L->pstar];
    T *D = (T *)L->pop[L->pstar ^ 1];
    if (a.s.force)
        hipLaunchKernelGGL((k_pull<T, true>), dim3(blocks), dim3(kPullThreads), 0, L->stream, S,
                           D, a);
    else
        hipLaunchKernelGGL((k_pull<T, false>), dim3(blocks), dim3(kPullThreads), 0, L->stream,
                           S, D, a);
    HIPCHK(hipGetLastError());
    L->pstar ^= 1;
    L->last = a.s;
    L->Fx_lag = L->cfg.Fx;
    L->Fy_lag = L->cfg.Fy;
    L->steps++;
    return LBMX_OK;
}

// A launch with no body force in its steps or in its first step's lagged moments
// runs the kernels' force-free instantiation (the same populations bit for bit).
template <typename A>
static bool no_force(const A &a)
{
    return !a.force && a.Fx == 0.0 && a.Fy == 0.0 && a.Fx_lag == 0.0 && a.Fy_lag == 0.0;
}

// Two gather-form steps in one launch (k_pull2); f*_{s+1} stays in LDS and only
// f*_{s+2} is stored.
template <typename T, bool W, int TY>
static void launch_pull2(lbmx_lattice *L, const T *S, T *D, const Pull2Args &a)
{
    const unsigned blocks = (unsigned)a.tiles_x * (unsigned)((L->Ny + TY - 1) / TY);
    if (no_force(a))
        hipLaunchKernelGGL((k_pull2<T, W, TY, true>), dim3(blocks), dim3(kPull2Threads), 0,
                           L->stream, S, D, a);
    else
        hipLaunchKernelGGL((k_pull2<T, W, TY>), dim3(blocks), dim3(kPull2Threads), 0, L->stream,
                           S, D, a);
}

template <typename T, bool W, int TY>
static void launch_pull3(lbmx_lattice *L, const T *S, T *D, const Pull2Args &a)
{
    const unsigned blocks = (unsigned)a.tiles_x * (unsigned)((L->Ny + TY - 1) / TY);
    if (no_force(a))
        hipLaunchKernelGGL((k_pull3<T, W, TY, true>), dim3(blocks), dim3(kPull2Threads), 0,
                           L->stream, S, D, a);
    else
        hipLaunchKernelGGL((k_pull3<T, W, TY>), dim3(blocks), dim3(kPull2Threads), 0, L->stream,
                           S, D, a);
}

// ---- row streaming (k_stream): D steps per launch ---------------------------------
// Depth, prefetch, ring and rows per workgroup were measured on the periodic
// Taylor-Green lattice at 1024^2 .. 16384^2: FP64 5 steps on large lattices (slim
// three-row rings, LDS 35.1 KB, four workgroups per CU) and 4 below, FP32 storage
// 8 steps with FP32 collision (two workgroups of 8 wavefronts) and 4 with FP64
// collision.
constexpr int kStreamLW = 64;

// Output columns per strip: the LW - 2(D - 1) a workgroup's last level holds,
// rounded down to whole 32-byte sectors of a row, so that neighbouring strips
// never write into one sector (FP32 D = 8: 50 -> 48 columns, D = 4: 58 -> 56;
// FP64 D = 4: 58 -> 56, D = 5: 56).  GLUPS full vs aligned: FP32 2048^2 199.5
// vs 209.4, 16384^2 279.2 vs 279.2; FP64 1536^2 100.0 vs 105.7.
static int stream_tx(int depth, int esize)
{
    const int full = kStreamLW - 2 * (depth - 1);
    const int q = 32 / esize;
    return full / q * q;
}

// Rows per workgroup: about 4096 workgroups over the lattice, as a power of two
// between 32 and 128 at depth 4, and between 128 (the depth-8 pipeline takes 21
// iterations to fill) and 512 at depth 8.  Depth 5 (FP64): 256 rows where that
// still gives 4096 workgroups, else 128 (GLUPS 128 / 256 / 512 rows: 16384^2
// 125.8 / 127.7-132.1 / 128.5, 8192^2 126.6 / 129.3, 4096^2 124.6 / 111.5).
static int stream_rows_base(int Nx, int Ny, int depth, int esize)
{
    const int tx = stream_tx(depth, esize);
    const double strips = (double)((Nx + tx - 1) / tx);
    if (depth == 5) return std::min(strips * ((Ny + 255) / 256) >= 4096 ? 256 : 128, Ny);
    const double want = strips * Ny / 4096.0;
    const int lo = depth <= 4 ? 32 : 128, hi = depth <= 4 ? 128 : 512;
    int rows = lo;
    while (rows < hi && want >= rows * 1.41421356) rows *= 2;
    return std::min(rows, Ny);
}

// Workgroups at the base row count (the launch thresholds below were measured
// with it).
static long long stream_wgs_of(int Nx, int Ny, int depth, int esize)
{
    const int tx = stream_tx(depth, esize);
    const int rows = stream_rows_base(Nx, Ny, depth, esize);
    return (long long)((Nx + tx - 1) / tx) * ((Ny + rows - 1) / rows);
}

// Lattices that give fewer than 768 workgroups of 32 rows at depth 4 run 16-row
// segments instead, which fill the CUs better despite the longer pipeline fill
// per row; GLUPS 32 vs 16 rows: FP64 768^2 54.4 vs 66.5, 1024^2 75.4 vs 77.1,
// 1280^2 (920 workgroups) 89.5 vs 85.0; FP32 1024^2 130.5 vs 140.9, 1280^2 167.2
// vs 155.5; FP32 + FP64 collision 1024^2 87.1 vs 92.8, 1280^2 111.0 vs 97.5.
//
// A grid a little larger than the workgroups the CUs hold at once (1,024 at four
// per CU on 256 CUs) runs its last partial round after the first: where 16-row
// segments give up to 1.5 times that many, the segments are lengthened until the
// grid fits one round (1024^2: 19 rows, 972 workgroups; GLUPS 16 vs 19 rows FP64
// 74.4 vs 83.2, FP32 139.0 vs 146.8, FP32 + FP64 collision 81.5 vs 96.6).
constexpr long long kStreamSlots = 1024;
static int stream_rows_of(int Nx, int Ny, int depth, int esize)
{
    const int rows = stream_rows_base(Nx, Ny, depth, esize);
    if (depth == 4 && rows == 32 && stream_wgs_of(Nx, Ny, depth, esize) < 768) {
        const int tx = stream_tx(depth, esize);
        const long long strips = (Nx + tx - 1) / tx;
        const long long wgs = strips * ((Ny + 15) / 16);
        if (wgs > kStreamSlots && wgs * 2 <= kStreamSlots * 3 && strips <= kStreamSlots) {
            const long long segs = kStreamSlots / strips;
            return (int)std::min<long long>(Ny, (Ny + segs - 1) / segs);
        }
        return 16;
    }
    return rows;
}
static int stream_rows(const lbmx_lattice *L, int depth)
{
    static const int rows_env = getenv("LBMX_STREAM_ROWS") ? atoi(getenv("LBMX_STREAM_ROWS")) : 0;
    if (rows_env > 0) return std::min(rows_env, L->Ny);
    return stream_rows_of(L->Nx, L->Ny, depth, (int)esize_of(L->cfg.precision));
}
static long long stream_threshold(int precision, int depth)
{
    return precision == LBMX_F32_WIDE ? 400 : (depth == 8 ? 320 : 200);
}

// Steps per streaming launch: 4; 8 with FP32 storage and FP32 collision from 6 M
// cells on, and whenever the caller forces streaming.  The four-step kernel is the
// faster one while the two FP32 population buffers mostly fit the 256 MB Infinity
// Cache, GLUPS four vs eight steps: 1408^2 163 vs 127, 1536^2 175 vs 151, 2048^2
// 192 vs 173; 3072^2 168 vs 205, 4096^2 153 vs 206, 8192^2 154 vs 218.
static long long stream_wgs_of(int Nx, int Ny, int depth, int esize);
// FP64 streams five steps per launch where the five-level grid has 512 workgroups
// of 128 rows (from about 2048 x 1800), and always when streaming is forced; four
// below.  GLUPS four vs five steps: 2048^2 100.4 vs 108.2, 4096^2 ~104 vs 124.6,
// 8192^2 98-112 vs 129.3, 16384^2 101-103 vs 127.7-132.1 (six steps: 119.9, three
// workgroups per CU).
static int stream_depth_of(int Nx, int Ny, int precision, int mode)
{
    if (precision == LBMX_F64)
        return (mode == LBMX_LAUNCH_STREAM || stream_wgs_of(Nx, Ny, 5, 8) >= 512) ? 5 : 4;
    if (precision != LBMX_F32_SHIFTED) return 4;
    if (mode == LBMX_LAUNCH_STREAM || (long long)Nx * Ny >= 6000000) return 8;
    return 4;
}
static int stream_depth(const lbmx_lattice *L)
{
    return stream_depth_of(L->Nx, L->Ny, L->cfg.precision, L->launch_mode);
}

// k_stream beats the three-step tiles once it has enough workgroups (counted at
// the base row count): 200 at four steps (FP64, FP32), 320 at eight, 400 with
// FP32 storage and FP64 collision, whose 16-row tiles stay strong longer.
// Measured, GLUPS tiles vs stream (base workgroups): FP64 512^2 56.3 vs 38.8
// (144), 576^2 47.4 vs 48.6 (180), 640^2 49.9 vs 59.9 (240), 704^2 54.5 vs 58.6
// (286); FP32 four steps 640^2 81.5 vs 100.8 (240), 768^2 100.0 vs 114.8 (336);
// FP32 eight steps 1280^2 121.9 vs 105.6 (260), 1536^2 134.3 vs 149.7 (372);
// FP32 + FP64 collision 704^2 76.0 vs 66.7 (286), 832^2 87.2 vs 78.9 (390),
// 896^2 72.3 vs 85.8 (448).  Smaller lattices run the tile kernels.
static bool stream_fits_of(int Nx, int Ny, int precision)
{
    const int depth = stream_depth_of(Nx, Ny, precision, LBMX_LAUNCH_AUTO);
    return stream_wgs_of(Nx, Ny, depth, (int)esize_of(precision)) >=
           stream_threshold(precision, depth);
}
static bool stream_fits(const lbmx_lattice *L)
{
    return stream_fits_of(L->Nx, L->Ny, L->cfg.precision);
}

static int tile_depth()
{
    static int depth_env = -1;
    if (depth_env < 0)
        depth_env = getenv("LBMX_TEMPORAL_DEPTH") ? atoi(getenv("LBMX_TEMPORAL_DEPTH")) : 0;
    return depth_env == 2 ? 2 : 3;
}

// Rows per tile of the multi-step tile kernels (kPull2TX columns each).  Tile
// heights measured at 16384^2 (ms per step): FP32 8 (1.70; 4: 1.69, 16: 1.70),
// FP32 storage + FP64 collision 16 (2.38; 8: 2.57), FP64 8.  The three-step tiles
// serve the lattices too small for row streaming, where more, shorter tiles pay
// off: the tallest of 8, 4, 2 rows that still gives 1024 workgroups (256^2: 2
// rows, FP64 19.6 vs 17.8 GLUPS with 8, FP32 26.4 vs 21.6; 512^2: 4 rows, FP32 77
// vs 72; 768^2 and up: 8 rows).
static int tile_rows_of(int Nx, int Ny, int precision, int depth)
{
    if (precision == LBMX_F32_WIDE) return 16;
    if (depth == 2) return 8;
    const long long tiles_x = (Nx + kPull2TX - 1) / kPull2TX;
    auto blocks = [&](int ty) { return tiles_x * ((Ny + ty - 1) / ty); };
    return blocks(8) >= 1024 ? 8 : (blocks(4) >= 1024 ? 4 : 2);
}

extern "C" int lbmx_launch_plan(int Nx, int Ny, int precision, int mode, int *steps_per_launch,
                                int *columns, int *rows)
{
    if (!steps_per_launch || !columns || !rows) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (Nx < 3 || Ny < 3) return lbmx_fail(LBMX_ERR_ARG, "lattice %dx%d too small", Nx, Ny);
    if (precision < LBMX_F64 || precision > LBMX_F32_WIDE)
        return lbmx_fail(LBMX_ERR_ARG, "unknown precision %d", precision);
    if (mode < LBMX_LAUNCH_AUTO || mode > LBMX_LAUNCH_STREAM)
        return lbmx_fail(LBMX_ERR_ARG, "unknown launch mode %d", mode);
    *columns = *rows = 0;
    if (mode == LBMX_LAUNCH_SINGLE) {
        *steps_per_launch = 1;
        return LBMX_OK;
    }
    const int sd = stream_depth_of(Nx, Ny, precision, mode);
    if (mode == LBMX_LAUNCH_STREAM || (mode == LBMX_LAUNCH_AUTO && stream_fits_of(Nx, Ny, precision))) {
        *steps_per_launch = sd;
        *columns = stream_tx(sd, (int)esize_of(precision));
        *rows = stream_rows_of(Nx, Ny, sd, (int)esize_of(precision));
        return LBMX_OK;
    }
    *steps_per_launch = tile_depth();
    if (mode == LBMX_LAUNCH_TILES || mode == LBMX_LAUNCH_AUTO) {
        *columns = kPull2TX;
        *rows = tile_rows_of(Nx, Ny, precision, *steps_per_launch);
    }
    return LBMX_OK;
}

template <typename T, bool W, int D, int P, int R>
static void launch_stream(lbmx_lattice *L, const T *S, T *Dst, StreamArgs &a)
{
    a.tx = stream_tx(D, sizeof(T));
    a.strips = (L->Nx + a.tx - 1) / a.tx;
    a.seg_rows = stream_rows(L, D);
    a.segs = (L->Ny + a.seg_rows - 1) / a.seg_rows;
    const unsigned blocks = ((unsigned)a.strips * (unsigned)a.segs + 7) / 8 * 8;
    if (no_force(a))
        hipLaunchKernelGGL((k_stream<T, W, D, kStreamLW, P, R, true>), dim3(blocks),
                           dim3(kStreamLW * D), 0, L->stream, S, Dst, a);
    else
        hipLaunchKernelGGL((k_stream<T, W, D, kStreamLW, P, R>), dim3(blocks),
                           dim3(kStreamLW * D), 0, L->stream, S, Dst, a);
}

template <typename T, bool W, int D, int P>
static void launch_wave(lbmx_lattice *L, const T *S, T *Dst, StreamArgs &a)
{
    a.tx = stream_tx(D, sizeof(T));
    a.strips = (L->Nx + a.tx - 1) / a.tx;
    a.seg_rows = stream_rows(L, D);
    a.segs = (L->Ny + a.seg_rows - 1) / a.seg_rows;
    const unsigned work = (unsigned)a.strips * (unsigned)a.segs;
    const unsigned blocks = ((work + kWaveWaves - 1) / kWaveWaves + 7) / 8 * 8;
    if (no_force(a))
        hipLaunchKernelGGL((k_wave<T, W, D, P, true>), dim3(blocks), dim3(64 * kWaveWaves), 0,
                           L->stream, S, Dst, a);
    else
        hipLaunchKernelGGL((k_wave<T, W, D, P>), dim3(blocks), dim3(64 * kWaveWaves), 0,
                           L->stream, S, Dst, a);
}

// `depth` gather-form steps in one launch (stream_depth(L), or 4 for four to seven
// steps left of a run at five or eight); as pull_step2 otherwise.
template <typename T>
static int pull_stream(lbmx_lattice *L, int depth)
{
    const Geo g = lbmx_geo(L);
    StepFlags s = lbmx_flags(L);
    StreamArgs a;
    a.g = g;
    a.force = s.force;
    a.Fx = s.Fx;
    a.Fy = s.Fy;
    a.Fx_lag = s.Fx_lag;
    a.Fy_lag = s.Fy_lag;
    const T *S = (const T *)L->pop[L->pstar];
    T *D = (T *)L->pop[L->pstar ^ 1];
    // FP64: slim three-row rings (five levels: 35.1 KB of LDS, four workgroups per
    // CU; the force-free kernel holds 80 VGPRs) with three rows of loads in flight
    // (two: 131.3, four: 121.4 GLUPS against 132.1 at 16384^2); FP32 storage with
    // FP32 collision: four-row rings.  FP32 storage with FP64 collision runs the
    // four levels in one wavefront's registers (k_wave): 160.8 vs 146.5 GLUPS at
    // 16384^2 (FP64 102.9 vs 102.8 at four steps, FP32 collision at eight steps
    // 207.6 vs 217.6, so those keep k_stream)
    if constexpr (sizeof(T) == 8) {
        if (depth == 5)
            launch_stream<T, false, 5, 3, 3>(L, S, D, a);
        else
            launch_stream<T, false, 4, 3, 3>(L, S, D, a);
    }
    else if (g.wide)
        launch_wave<T, true, 4, 2>(L, S, D, a);
    else if (depth == 8)
        launch_stream<T, false, 8, 6, 4>(L, S, D, a);
    else
        launch_stream<T, false, 4, 4, 4>(L, S, D, a);
    HIPCHK(hipGetLastError());
    L->pstar ^= 1;
    s.Fx_lag = L->cfg.Fx;
    s.Fy_lag = L->cfg.Fy;
    L->last = s;
    L->Fx_lag = L->cfg.Fx;
    L->Fy_lag = L->cfg.Fy;
    L->steps += depth;
    return LBMX_OK;
}

// Three gather-form steps in one launch (k_pull3); as pull_step2 otherwise.
template <typename T>
static int pull_step3(lbmx_lattice *L)
{
    const Geo g = lbmx_geo(L);
    StepFlags s = lbmx_flags(L);
    Pull2Args a;
    a.g = g;
    a.force = s.force;
    a.Fx = s.Fx;
    a.Fy = s.Fy;
    a.Fx_lag = s.Fx_lag;
    a.Fy_lag = s.Fy_lag;
    a.tiles_x = (L->Nx + kPull2TX - 1) / kPull2TX;
    const T *S = (const T *)L->pop[L->pstar];
    T *D = (T *)L->pop[L->pstar ^ 1];
    const int ty = tile_rows_of(L->Nx, L->Ny, L->cfg.precision, 3);
    bool done = false;
    if constexpr (sizeof(T) == 4) {
        if (g.wide) {
            launch_pull3<T, true, 16>(L, S, D, a);
            done = true;
        }
    }
    if (!done) {
        if (ty == 8)
            launch_pull3<T, false, 8>(L, S, D, a);
        else if (ty == 4)
            launch_pull3<T, false, 4>(L, S, D, a);
        else
            launch_pull3<T, false, 2>(L, S, D, a);
    }
    HIPCHK(hipGetLastError());
    L->pstar ^= 1;
    s.Fx_lag = L->cfg.Fx;
    s.Fy_lag = L->cfg.Fy;
    L->last = s;
    L->Fx_lag = L->cfg.Fx;
    L->Fy_lag = L->cfg.Fy;
    L->steps += 3;
    return LBMX_OK;
}

template <typename T>
static int pull_step2(lbmx_lattice *L)
{
    const Geo g = lbmx_geo(L);
    StepFlags s = lbmx_flags(L);
    Pull2Args a;
    a.g = g;
    a.force = s.force;
    a.Fx = s.Fx;
    a.Fy = s.Fy;
    a.Fx_lag = s.Fx_lag;
    a.Fy_lag = s.Fy_lag;
    a.tiles_x = (L->Nx + kPull2TX - 1) / kPull2TX;
    const T *S = (const T *)L->pop[L->pstar];
    T *D = (T *)L->pop[L->pstar ^ 1];
    // tile heights (LDS of a workgroup: 9 x 66 x (TY + 2) values, under 64 KB) as
    // measured fastest at 16384^2 among 2-24: FP64 8 (3.72 ms per step; 4: 3.82,
    // 10: 4.23), FP32 8 (1.91; 16: 1.94, 24: 2.25), FP32 storage + FP64
    // collision 16 (2.57; 8: 2.72); tile_rows_of(..., 2) reports them
    if constexpr (sizeof(T) == 8)
        launch_pull2<T, false, 8>(L, S, D, a);
    else if (g.wide)
        launch_pull2<T, true, 16>(L, S, D, a);
    else
        launch_pull2<T, false, 8>(L, S, D, a);
    HIPCHK(hipGetLastError());
    L->pstar ^= 1;
    s.Fx_lag = L->cfg.Fx;
    s.Fy_lag = L->cfg.Fy;
    L->last = s;
    L->Fx_lag = L->cfg.Fx;
    L->Fy_lag = L->cfg.Fy;
    L->steps += 2;
    return LBMX_OK;
}

int lbmx_ensure_push(lbmx_lattice *L)
{
    if (!L->pull) return LBMX_OK;
    // pop[pstar] holds f*_{k-1}, the last step's post-collision field: its gather
    // is the newest pre-collision state f_k, and it stays as the reference's f_new
    const Geo g = lbmx_geo(L);
    dim3 grid((L->Nx + 255) / 256, L->Ny);
    if (L->esize == 8)
        hipLaunchKernelGGL(k_gather<double>, grid, dim3(256), 0, L->stream,
                           (const double *)L->pop[L->pstar], (double *)L->pop[L->pstar ^ 1], g);
    else
        hipLaunchKernelGGL(k_gather<float>, grid, dim3(256), 0, L->stream,
                           (const float *)L->pop[L->pstar], (float *)L->pop[L->pstar ^ 1], g);
    HIPCHK(hipGetLastError());
    L->cur = L->pstar ^ 1;
    L->pull = 0;
    L->other_is_ref = 1;
    L->lag_arrays = 0;
    return LBMX_OK;
}

extern "C" int lbmx_set_launch(lbmx_lattice *L, int mode)
{
    if (!L) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (mode < LBMX_LAUNCH_AUTO || mode > LBMX_LAUNCH_STREAM)
        return lbmx_fail(LBMX_ERR_ARG, "unknown launch mode %d", mode);
    L->launch_mode = mode;
    return LBMX_OK;
}

extern "C" int lbmx_launch_depth(const lbmx_lattice *L)
{
    if (!L) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (!pull_eligible(L)) {
        lbmx_lattice *M = const_cast<lbmx_lattice *>(L);   // caches the eligibility check
        const bool on = L->launch_mode == LBMX_LAUNCH_STREAM ||
                        (L->launch_mode == LBMX_LAUNCH_AUTO && lbmx_chan_fits(L));
        return on && lbmx_chan_eligible(M) ? lbmx_chan_depth() : 1;
    }
    int d = 1, c = 0, r = 0;
    const int rc = lbmx_launch_plan(L->Nx, L->Ny, L->cfg.precision, L->launch_mode, &d, &c, &r);
    return rc ? rc : d;
}

extern "C" int lbmx_step(lbmx_lattice *L, int nsteps)
{
    if (!L) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (!L->has_state) return lbmx_fail(LBMX_ERR_STATE, "no state uploaded");
    for (size_t b = 0; b < L->vals_set.size(); b++)
        if (lbmx_bc_nvals(L->types[b]) && !L->vals_set[b])
            return lbmx_fail(LBMX_ERR_STATE, "boundary %zu has no values", b);
    if (nsteps <= 0) return LBMX_OK;
    const bool f64 = (L->esize == 8);
    const int64_t steps0 = L->steps;
    const int mode = L->launch_mode;
    if (!L->pull && nsteps >= 2 && pull_eligible(L)) {
        int rc = f64 ? enter_pull<double>(L) : enter_pull<float>(L);
        if (rc) return rc;
    }
    if (L->pull && !pull_eligible(L)) {
        int rc = lbmx_ensure_push(L);
        if (rc) return rc;
    }
    int i = L->steps - steps0;   // 1 when entering the gather form took the first step
    // bounded lattices (walls, bodies, regularized columns): several steps per
    // launch when row streaming is on for them, the call's last 1..depth steps
    // through the one-step path (which leaves both buffers as the reference's)
    if (!L->pull && (mode == LBMX_LAUNCH_STREAM || (mode == LBMX_LAUNCH_AUTO && lbmx_chan_fits(L))) &&
        nsteps > lbmx_chan_depth() && lbmx_chan_eligible(L)) {
        const int cd = lbmx_chan_depth();
        int rc = LBMX_OK;
        for (const int n = (nsteps - 1) / cd; rc == LBMX_OK && i < n * cd; i += cd)
            rc = lbmx_chan_run(L);
        const int rf = lbmx_chan_flush(L);   // the last launch's forces
        if (rc || rf) return rc ? rc : rf;
    }
    // tile depth: 3 steps per launch (the tiles serve the lattices too small for
    // row streaming, where three beat two in FP64 as in FP32: 256^2 3.69 vs 3.97
    // us per step, 512^2 4.66 vs 5.05); LBMX_TEMPORAL_DEPTH=2 or 3 overrides
    const int depth = tile_depth();
    if (L->pull && mode != LBMX_LAUNCH_SINGLE) {
        // multi-step launches while at least as many steps remain as they take; the
        // buffer a launch writes is the post-collision field of its last step, all a
        // following run or the way back to the reference's buffers needs
        const int sd = stream_depth(L);
        if (mode == LBMX_LAUNCH_STREAM || (mode == LBMX_LAUNCH_AUTO && stream_fits(L))) {
            for (; nsteps - i >= sd; i += sd) {
                int rc = f64 ? pull_stream<double>(L, sd) : pull_stream<float>(L, sd);
                if (rc) return rc;
            }
            // four to seven steps left of a run at five (FP64) or eight (FP32): one
            // four-level launch first rather than three-step tiles and single steps
            // (FP64 16384^2: 10.5 ms for four steps vs 10.4 + 7.1 ms)
            if (sd > 4 && nsteps - i >= 4) {
                int rc = f64 ? pull_stream<double>(L, 4) : pull_stream<float>(L, 4);
                if (rc) return rc;
                i += 4;
            }
        }
        if (depth >= 3)
            for (; nsteps - i >= 3; i += 3) {
                int rc = f64 ? pull_step3<double>(L) : pull_step3<float>(L);
                if (rc) return rc;
            }
        for (; nsteps - i >= 2; i += 2) {
            int rc = f64 ? pull_step2<double>(L) : pull_step2<float>(L);
            if (rc) return rc;
        }
    }
    for (; i < nsteps; i++) {
        int rc;
        if (L->pull)
            rc = f64 ? pull_step<double>(L) : pull_step<float>(L);
        else
            rc = f64 ? launch_step<double>(L) : launch_step<float>(L);
        if (rc) return rc;
    }
    return LBMX_OK;
}

// ---------------------------------------------------------------------------
// time-domain impedance boundary controller
// ---------------------------------------------------------------------------
static void tdibc_free(lbmx_lattice *L)
{
    hipFree(L->tdibc.st);
    hipFree(L->tdibc.last);
    hipFree(L->tdibc.active);
    hipFree(L->tdibc.hist);
    L->tdibc.st = nullptr;
    L->tdibc.last = nullptr;
    L->tdibc.active = nullptr;
    L->tdibc.hist = nullptr;
    L->tdibc.on = 0;
}

extern "C" int lbmx_set_tdibc(lbmx_lattice *L, const lbmx_tdibc *c)
{
    if (!L || !c) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (c->bc < 0 || c->bc >= (int)L->types.size())
        return lbmx_fail(LBMX_ERR_ARG, "tdibc: boundary %d not in the table", c->bc);
    const int t = L->types[c->bc];
    if (c->which < 1 || c->which > lbmx_bc_nvals(t))
        return lbmx_fail(LBMX_ERR_ARG, "tdibc: boundary type %d has no val%d", t, c->which);
    const int col = lbmx_bc_on_column(t);
    if ((c->probe_is_column != 0) != (col != 0))
        return lbmx_fail(LBMX_ERR_ARG, "tdibc: probe line must run along the boundary");
    const int nedge = col ? L->Ny : L->Nx;
    const int span = col ? L->Nx : L->Ny;
    if (c->probe_index < 0 || c->probe_index >= span)
        return lbmx_fail(LBMX_ERR_ARG, "tdibc: probe index %d out of range", c->probe_index);
    if (c->capacity < 0) return lbmx_fail(LBMX_ERR_ARG, "tdibc: negative capacity");
    HIPCHK(hipStreamSynchronize(L->stream));
    if (int prc = lbmx_ensure_push(L)) return prc;
    tdibc_free(L);
    TdibcParams p{};
    p.which = c->which;
    p.probe_is_column = c->probe_is_column ? 1 : 0;
    p.probe_index = c->probe_index;
    p.per_node = c->per_node ? 1 : 0;
    p.record_node = c->record_node;
    p.capacity = c->capacity;
    p.nedge = nedge;
    p.rho0 = c->rho0;
    p.cs = c->cs;
    p.cp = c->cp;
    p.dt = c->dt;
    p.cu = c->cu;
    p.gain = c->gain;
    p.Zinf = c->Zinf;
    for (int i = 0; i < 2; i++) {
        p.A[i] = c->A[i];
        p.lambdas[i] = c->lambdas[i];
        p.B[i] = c->B[i];
        p.C[i] = c->C[i];
        p.alpha[i] = c->alpha[i];
        p.beta[i] = c->beta[i];
    }
    const int nst = p.per_node ? nedge : 1;
    HIPCHK(hipMalloc((void **)&L->tdibc.st, nst * sizeof(TdibcState)));
    HIPCHK(hipMemset(L->tdibc.st, 0, nst * sizeof(TdibcState)));
    HIPCHK(hipMalloc((void **)&L->tdibc.last, 2 * sizeof(double)));
    HIPCHK(hipMemset(L->tdibc.last, 0, 2 * sizeof(double)));
    if (c->active) {
        HIPCHK(hipMalloc((void **)&L->tdibc.active, nedge * sizeof(int)));
        HIPCHK(hipMemcpy(L->tdibc.active, c->active, nedge * sizeof(int),
                         hipMemcpyHostToDevice));
    }
    if (p.capacity > 0) {
        HIPCHK(hipMalloc((void **)&L->tdibc.hist, (size_t)p.capacity * 3 * sizeof(double)));
        HIPCHK(hipMemset(L->tdibc.hist, 0, (size_t)p.capacity * 3 * sizeof(double)));
    }
    L->tdibc.p = p;
    L->tdibc.bc = c->bc;
    L->tdibc.step = 0;
    L->tdibc.on = 1;
    // the controller provides this boundary's driven values from now on
    L->vals_set[c->bc] = 1;
    return LBMX_OK;
}

extern "C" int lbmx_tdibc_history(lbmx_lattice *L, double *rec, int rows)
{
    if (!L || !rec) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (!L->tdibc.on) return lbmx_fail(LBMX_ERR_STATE, "no tdibc controller");
    long long n = L->tdibc.step < L->tdibc.p.capacity ? L->tdibc.step : L->tdibc.p.capacity;
    if (n > rows) n = rows;
    if (n > 0)
        HIPCHK(hipMemcpyAsync(rec, L->tdibc.hist, (size_t)n * 3 * sizeof(double),
                              hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    return (int)n;
}

template <typename T>
static int launch_tdibc(lbmx_lattice *L, const T *S, const Geo &g, const StepFlags &s,
                        const Fields &F)
{
    TdibcArgs a;
    a.g = g;
    a.s = s;
    a.F = F;
    a.p = L->tdibc.p;
    a.st = L->tdibc.st;
    a.last = L->tdibc.last;
    a.active = L->tdibc.active;
    a.val = (L->tdibc.p.which == 1) ? L->val1[L->tdibc.bc] : L->val2[L->tdibc.bc];
    a.hist = L->tdibc.hist;
    a.step = L->tdibc.step;
    if (a.p.per_node)
        hipLaunchKernelGGL(k_tdibc_nodes<T>, dim3((a.p.nedge + 255) / 256), dim3(256), 0,
                           L->stream, S, a);
    else
        hipLaunchKernelGGL(k_tdibc_lumped<T>, dim3(1), dim3(kTdibcThreads), 0, L->stream, S, a);
    HIPCHK(hipGetLastError());
    L->tdibc.step++;
    return LBMX_OK;
}

extern "C" int lbmx_set_probes(lbmx_lattice *L, const int *x, const int *y, int n, int capacity)
{
    if (!L || n < 0 || capacity < 0 || (n > 0 && (!x || !y)))
        return lbmx_fail(LBMX_ERR_ARG, "bad probe list");
    if (int prc = lbmx_ensure_push(L)) return prc;
    std::vector<long long> nodes(n);
    for (int i = 0; i < n; i++) {
        if (x[i] < 0 || x[i] >= L->Nx || y[i] < 0 || y[i] >= L->Ny)
            return lbmx_fail(LBMX_ERR_ARG, "probe %d at (%d, %d) outside the lattice", i, x[i],
                             y[i]);
        nodes[i] = (long long)y[i] * L->pitch + x[i];
    }
    HIPCHK(hipStreamSynchronize(L->stream));
    hipFree(L->probes.nodes);
    hipFree(L->probes.hist);
    L->probes.nodes = nullptr;
    L->probes.hist = nullptr;
    L->probes.n = 0;
    if (n == 0 || capacity == 0) return LBMX_OK;
    HIPCHK(hipMalloc((void **)&L->probes.nodes, n * sizeof(long long)));
    HIPCHK(hipMemcpy(L->probes.nodes, nodes.data(), n * sizeof(long long),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMalloc((void **)&L->probes.hist, (size_t)n * capacity * sizeof(double)));
    HIPCHK(hipMemset(L->probes.hist, 0, (size_t)n * capacity * sizeof(double)));
    L->probes.n = n;
    L->probes.capacity = capacity;
    L->probes.step = 0;
    return LBMX_OK;
}

extern "C" int lbmx_probe_history(lbmx_lattice *L, double *rec, int rows)
{
    if (!L || !rec) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (!L->probes.n) return lbmx_fail(LBMX_ERR_STATE, "no probes");
    long long m = L->probes.step < L->probes.capacity ? L->probes.step : L->probes.capacity;
    if (m > rows) m = rows;
    if (m > 0)
        HIPCHK(hipMemcpyAsync(rec, L->probes.hist, (size_t)m * L->probes.n * sizeof(double),
                              hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    return (int)m;
}

template <typename T>
static int launch_probes(lbmx_lattice *L, const T *S, const Geo &g, const StepFlags &s,
                         const Fields &F)
{
    ProbeArgs a;
    a.g = g;
    a.s = s;
    a.F = F;
    a.nodes = L->probes.nodes;
    a.n = L->probes.n;
    a.capacity = L->probes.capacity;
    a.hist = L->probes.hist;
    a.step = L->probes.step;
    hipLaunchKernelGGL(k_probes<T>, dim3((a.n + 255) / 256), dim3(256), 0, L->stream, S, a);
    HIPCHK(hipGetLastError());
    L->probes.step++;
    return LBMX_OK;
}

extern "C" int lbmx_time_steps(lbmx_lattice *L, int nsteps, double *ms)
{
    if (!L || !ms) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, L->stream));
    int rc = lbmx_step(L, nsteps);
    if (rc) {
        hipEventDestroy(a);
        hipEventDestroy(b);
        return rc;
    }
    HIPCHK(hipEventRecord(b, L->stream));
    HIPCHK(hipEventSynchronize(b));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    *ms = t;
    return LBMX_OK;
}

extern "C" int lbmx_forces(lbmx_lattice *L, double *Lift, double *Drag)
{
    if (!L) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    HIPCHK(hipMemcpyAsync(L->force_host, L->force_dev, kPartial * sizeof(double),
                          hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    if (Drag) *Drag = L->force_host[0];
    if (Lift) *Lift = L->force_host[1];
    return LBMX_OK;
}

extern "C" int lbmx_body_forces(lbmx_lattice *L, double *Lift, double *Drag)
{
    if (!L) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    HIPCHK(hipMemcpyAsync(L->force_host, L->force_dev, kPartial * sizeof(double),
                          hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    if (Drag) *Drag = L->force_host[2];
    if (Lift) *Lift = L->force_host[3];
    return LBMX_OK;
}

extern "C" int lbmx_set_force_body(lbmx_lattice *L, const int *body)
{
    if (!L) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (!L->has_geometry) return lbmx_fail(LBMX_ERR_STATE, "lbmx_set_geometry first");
    if (int prc = lbmx_ensure_push(L)) return prc;
    if (L->nlinks == 0) return LBMX_OK;
    HIPCHK(hipStreamSynchronize(L->stream));
    for (size_t l = 0; l < L->hlinks.size(); l++)
        L->hlinks[l].body = (body && body[L->link_target[l]]) ? 1 : 0;
    HIPCHK(hipMemcpy(L->links, L->hlinks.data(), L->hlinks.size() * sizeof(Link),
                     hipMemcpyHostToDevice));
    return LBMX_OK;
}

extern "C" int lbmx_set_force_record(lbmx_lattice *L, double dt, double cf, int capacity)
{
    if (!L || capacity < 0) return lbmx_fail(LBMX_ERR_ARG, "invalid argument");
    if (int prc = lbmx_ensure_push(L)) return prc;
    HIPCHK(hipStreamSynchronize(L->stream));
    if (L->frec) hipFree(L->frec);
    L->frec = nullptr;
    L->frec_cap = 0;
    L->frec_row = 0;
    if (capacity == 0) return LBMX_OK;
    if (hipMalloc((void **)&L->frec, (size_t)capacity * kRecordCols * sizeof(double)) !=
        hipSuccess) {
        L->frec = nullptr;
        return lbmx_fail(LBMX_ERR_NOMEM, "force record of %d rows", capacity);
    }
    L->frec_cap = capacity;
    L->frec_dt = dt;
    L->frec_cf = cf;
    return LBMX_OK;
}

extern "C" int lbmx_force_record(lbmx_lattice *L, double *rows, int capacity)
{
    if (!L || !rows || capacity < 0) return lbmx_fail(LBMX_ERR_ARG, "invalid argument");
    if (!L->frec) return 0;
    long long n = L->frec_row < L->frec_cap ? L->frec_row : L->frec_cap;
    if (n > capacity) n = capacity;
    HIPCHK(hipStreamSynchronize(L->stream));
    if (n > 0)
        HIPCHK(hipMemcpy2D(rows, 3 * sizeof(double), L->frec, kRecordCols * sizeof(double),
                           3 * sizeof(double), (size_t)n, hipMemcpyDeviceToHost));
    return (int)n;
}

extern "C" int lbmx_body_force_record(lbmx_lattice *L, double *rows, int capacity)
{
    if (!L || !rows || capacity < 0) return lbmx_fail(LBMX_ERR_ARG, "invalid argument");
    if (!L->frec) return 0;
    long long n = L->frec_row < L->frec_cap ? L->frec_row : L->frec_cap;
    if (n > capacity) n = capacity;
    HIPCHK(hipStreamSynchronize(L->stream));
    if (n > 0) {
        std::vector<double> all((size_t)n * kRecordCols);
        HIPCHK(hipMemcpy(all.data(), L->frec, all.size() * sizeof(double),
                         hipMemcpyDeviceToHost));
        for (long long r = 0; r < n; r++) {
            rows[3 * r + 0] = all[kRecordCols * r + 0];
            rows[3 * r + 1] = all[kRecordCols * r + 3];
            rows[3 * r + 2] = all[kRecordCols * r + 4];
        }
    }
    return (int)n;
}

int lbmx_launch_moments(lbmx_lattice *L)
{
    if (!L->Ro) {
        const size_t fld = (size_t)L->Ny * L->pitch * sizeof(double);
        if (hipMalloc((void **)&L->Ro, fld) != hipSuccess ||
            hipMalloc((void **)&L->Uo, fld) != hipSuccess ||
            hipMalloc((void **)&L->Vo, fld) != hipSuccess)
            return lbmx_fail(LBMX_ERR_NOMEM, "moment output arrays (%zu bytes)", 3 * fld);
        L->bytes += 3 * fld;
    }
    MomArgs m;
    m.g = lbmx_geo(L);
    m.F = Fields{L->Ro, L->Uo, L->Vo};
    m.Fx = L->Fx_lag;   // the force of the last step, as its compute_macroscopic_fields took it
    m.Fy = L->Fy_lag;
    dim3 grid((L->Nx + 255) / 256, L->Ny);
    if (L->esize == 8)
        hipLaunchKernelGGL(k_moments<double>, grid, dim3(256), 0, L->stream,
                           (const double *)L->pop[L->cur], m);
    else
        hipLaunchKernelGGL(k_moments<float>, grid, dim3(256), 0, L->stream,
                           (const float *)L->pop[L->cur], m);
    HIPCHK(hipGetLastError());
    return LBMX_OK;
}

extern "C" int lbmx_moments(lbmx_lattice *L, double *rho, double *u, double *v)
{
    if (!L) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (!L->has_state) return lbmx_fail(LBMX_ERR_STATE, "no state uploaded");
    if (int prc = lbmx_ensure_push(L)) return prc;
    // before the first step the arrays are the uploaded ones, as the reference's
    // would be; afterwards compute_macroscopic_fields' result of the last step
    const double *r = L->R, *uu = L->U, *vv = L->V;
    if (L->steps > 0) {
        int rc = lbmx_launch_moments(L);
        if (rc) return rc;
        r = L->Ro;
        uu = L->Uo;
        vv = L->Vo;
    }
    int rc = LBMX_OK;
    if (rho) rc = lbmx_get_field(L, rho, r);
    if (!rc && u) rc = lbmx_get_field(L, u, uu);
    if (!rc && v) rc = lbmx_get_field(L, v, vv);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(L->stream));
    return LBMX_OK;
}

// ---------------------------------------------------------------------------
// VTK output packed on the device (write_vtk_binary_2D.c:47-78 layout)
// ---------------------------------------------------------------------------
extern "C" void lbm_vtk_head(FILE *fd, int Nx, int Ny, double dx);
extern "C" void lbm_vtk_mid(FILE *fd);

// (float)(x * c) as big-endian bits.  A NaN (solid nodes) converts as the x86
// host writer converts it: sign and the top mantissa bits kept, quiet bit set.
__device__ __forceinline__ uint32_t vtk_be(double x, double c)
{
    uint32_t b;
    if (isnan(x)) {
        const unsigned long long d = (unsigned long long)__double_as_longlong(x);
        b = (uint32_t)((d >> 32) & 0x80000000ull) | 0x7fc00000u | (uint32_t)((d >> 29) & 0x3fffffull);
    } else {
        b = __float_as_uint((float)(x * c));
    }
    return __builtin_bswap32(b);
}

__global__ void __launch_bounds__(256) k_vtk_pack(const double *__restrict__ R,
                                                  const double *__restrict__ U,
                                                  const double *__restrict__ V, int Nx,
                                                  int pitch, long long nn, double cu,
                                                  double crho, uint32_t *__restrict__ out)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= Nx) return;
    const size_t s = (size_t)y * pitch + x;
    const size_t n = (size_t)y * Nx + x;
    out[3 * n + 0] = vtk_be(U[s], cu);
    out[3 * n + 1] = vtk_be(V[s], cu);
    out[3 * n + 2] = 0u;   // big-endian 0.0f
    out[3 * nn + n] = vtk_be(R[s], crho);
}

extern "C" int lbmx_write_vtk(lbmx_lattice *L, const char *filename, double dx, double cu,
                              double crho)
{
    if (!L || !filename) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (!L->has_state) return lbmx_fail(LBMX_ERR_STATE, "no state uploaded");
    if (int prc = lbmx_ensure_push(L)) return prc;
    const double *r = L->R, *u = L->U, *v = L->V;
    if (L->steps > 0) {
        int rc = lbmx_launch_moments(L);
        if (rc) return rc;
        r = L->Ro;
        u = L->Uo;
        v = L->Vo;
    }
    const long long nn = (long long)L->Nx * L->Ny;
    const size_t chunk = (size_t)32 << 20;   // bytes per pinned host transfer
    if (!L->vtk_dev) {
        if (hipMalloc((void **)&L->vtk_dev, (size_t)nn * 16) != hipSuccess)
            return lbmx_fail(LBMX_ERR_NOMEM, "VTK pack buffer (%lld bytes)", nn * 16);
        L->bytes += nn * 16;
    }
    if (!L->vtk_host) HIPCHK(hipHostMalloc((void **)&L->vtk_host, chunk));
    dim3 grid((L->Nx + 255) / 256, L->Ny);
    hipLaunchKernelGGL(k_vtk_pack, grid, dim3(256), 0, L->stream, r, u, v, L->Nx, L->pitch, nn,
                       cu, crho, L->vtk_dev);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(L->stream));

    FILE *fd = fopen(filename, "wb");
    if (!fd) return lbmx_fail(LBMX_ERR_ARG, "cannot open %s for writing", filename);
    int rc = LBMX_OK;
    // velocity block (12 B per node), the text between, the density block (4 B)
    auto emit = [&](const uint32_t *dev, size_t bytes) -> int {
        for (size_t off = 0; off < bytes; off += chunk) {
            const size_t len = std::min(chunk, bytes - off);
            HIPCHK(hipMemcpyAsync(L->vtk_host, (const char *)dev + off, len,
                                  hipMemcpyDeviceToHost, L->stream));
            HIPCHK(hipStreamSynchronize(L->stream));
            if (fwrite(L->vtk_host, 1, len, fd) != len)
                return lbmx_fail(LBMX_ERR_ARG, "short write to %s", filename);
        }
        return LBMX_OK;
    };
    lbm_vtk_head(fd, L->Nx, L->Ny, dx);
    rc = emit(L->vtk_dev, (size_t)nn * 12);
    if (!rc) {
        lbm_vtk_mid(fd);
        rc = emit(L->vtk_dev + 3 * nn, (size_t)nn * 4);
    }
    if (fclose(fd) != 0 && !rc) rc = lbmx_fail(LBMX_ERR_ARG, "cannot close %s", filename);
    return rc;
}

extern "C" int lbmx_totals(lbmx_lattice *L, double *mass, double *energy)
{
    if (!L || !mass || !energy) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (!L->has_state) return lbmx_fail(LBMX_ERR_STATE, "no state uploaded");
    if (int prc = lbmx_ensure_push(L)) return prc;
    TotalsArgs a{};
    a.R = L->R;
    a.U = L->U;
    a.V = L->V;
    if (L->steps > 0) {
        if (int rc = lbmx_launch_moments(L)) return rc;
        a.R = L->Ro;
        a.U = L->Uo;
        a.V = L->Vo;
    }
    a.Nx = L->Nx;
    a.Ny = L->Ny;
    a.pitch = L->pitch;
    a.cls = L->geom ? L->cls : nullptr;
    if (!L->tot_dev) {
        HIPCHK(hipMalloc((void **)&L->tot_dev, sizeof(double) * (2 * kTotMaxBlocks + 2)));
        L->bytes += sizeof(double) * (2 * kTotMaxBlocks + 2);
    }
    double *out = L->tot_dev + 2 * kTotMaxBlocks;
    HIPCHK(launch_totals(a, L->tot_dev, out, L->stream));
    double t[2];
    HIPCHK(hipMemcpyAsync(t, out, sizeof t, hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    *mass = t[0];
    *energy = t[1];
    return LBMX_OK;
}

extern "C" int lbmx_synchronize(lbmx_lattice *L)
{
    if (!L) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(L->stream));
    return LBMX_OK;
}

extern "C" int64_t lbmx_steps_done(const lbmx_lattice *L) { return L ? L->steps : 0; }
extern "C" int64_t lbmx_device_bytes(const lbmx_lattice *L) { return L ? L->bytes : 0; }
