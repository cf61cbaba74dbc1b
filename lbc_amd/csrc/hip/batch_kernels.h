// batch_kernels.h -- gfx950 kernels that advance many independent lattices in one
// persistent launch per time step.
//
// Storage of one lattice (one arena block): two population buffers of nine
// dense planes, node n = y*Nx + x, plane stride `plane` elements (a multiple of
// 64, so every plane starts 256-byte aligned).  The buffers hold POST-collision
// populations f* (gather form): a step reads f*_s from buf[s & 1], forms the
// pre-collision populations of the next step by pulling from the neighbours
// (streaming.c:21-42), applies the case's boundary closures, takes the moments,
// collides (collision.c:24-37) and writes f*_{s+1} into buf[(s + 1) & 1].
//
// Reference semantics kept exactly (src/core/main_lbm.c:31-61):
//   * solid sources give half-way bounce-back, f(x, k) = f*(x, opp k)
//     (bounce_back.c:43-46); solid nodes are never written.
//   * sources outside a non-periodic domain leave the entry stale
//     (streaming.c:36-37): it keeps f*_{s-1}(x, k), which is exactly what the
//     destination buffer holds before this step overwrites it.
//   * collision uses the moments left by the previous step, i.e. the moments of
//     the final (post-boundary) populations it collides; the first collision uses
//     the initial moment arrays (done by k_batch_init).
//   * boundary closures that read lagged moments (the regularized wall) get them
//     from the neighbour's previous step, kept in a small per-lattice array.
// The FP64 arithmetic is d2q9.h's (reference rounding); FP32-shifted storage uses
// step_kernels.h's moments_f32/collide_f32 and decodes to FP64 for the boundary
// closures, exactly as the single-lattice kernels do.
#pragma once
#include "step_kernels.h"
#include "tdibc_kernels.h"

namespace lbmx {

enum : unsigned {
    BF_PER_X = 1u << 0,
    BF_PER_Y = 1u << 1,
    BF_WALL_BOTTOM = 1u << 2,
    BF_WALL_TOP = 1u << 3,
    BF_WALL_LEFT = 1u << 4,
    BF_WALL_RIGHT = 1u << 5,
    BF_TOP_VELOCITY = 1u << 6,
    BF_CONVECTIVE = 1u << 7,
    BF_LINER = 1u << 8,
    BF_FORCE = 1u << 9,
    BF_MOMENTS = 1u << 10,
};

// Per-lattice descriptor, read with scalar loads (one lattice per tile).
struct BatchLat {
    void *buf[2];            // population buffers (T[9][plane])
    const double *om_col;    // relaxation rate per column
    double *out[3];          // rho, u, v at completion (BF_MOMENTS)
    double *saved[2];        // liner: lagged (rho, u, v) of row 1 over the opening, by parity
    double *state;           // liner: phi[2] psi[2] chi[2] rho_lag (8 doubles) per opening node
    double *probe_hist;      // (steps + 1) x nprobe
    long long start;         // batch epoch at which this lattice takes its step 0
    unsigned plane;
    int Nx, Ny;
    unsigned flags;
    int gap_x0, gap_x1;
    int steps;
    int nprobe;
    int probe_n[4];
    double rho0, Fx, Fy, U_top, V_top;
    TdibcParams tp;
};

struct BatchCtl {
    long long epoch;         // steps launched so far (k_batch_advance)
    int ntiles;
    int pad;
};

struct Tile {
    int lat;                 // descriptor slot
    int start;               // first node
};

// Initial state sources of k_batch_init.
struct BatchInit {
    int kind;                // LBMX_INIT_PULSE / _TGV / _FIELDS
    const double *prof;      // pulse: rho per column
    const double *tx;        // TGV: sin(kx x), cos(kx x), cos(2 kx x) per column (3 x Nx)
    const double *ty;        // TGV: sin(ky y), cos(ky y), cos(2 ky y) per row (3 x Ny)
    const double *fields;    // rho, u, v planes (3 x Nx*Ny)
    double u0, r;            // TGV amplitude and kx/ky
};

constexpr int kBatchThreads = 256;

__device__ __forceinline__ bool b_solid(const BatchLat &L, int x, int y)
{
    const unsigned F = L.flags;
    return ((F & BF_WALL_BOTTOM) && y == 0 && (x < L.gap_x0 || x > L.gap_x1)) ||
           ((F & BF_WALL_TOP) && y == L.Ny - 1) || ((F & BF_WALL_LEFT) && x == 0) ||
           ((F & BF_WALL_RIGHT) && x == L.Nx - 1);
}

// stored value <-> population (FP32-shifted stores f - w*rho0, as pop_get/pop_put)
__device__ __forceinline__ double b_dec(double v, int, double) { return v; }
__device__ __forceinline__ double b_dec(float v, int k, double rho0)
{
    return (double)v + D2Q9::w[k] * rho0;
}
template <typename T>
__device__ __forceinline__ T b_enc(double v, int k, double rho0)
{
    if constexpr (sizeof(T) == 8)
        return v;
    else
        return (float)(v - D2Q9::w[k] * rho0);
}

// Pre-collision populations of node (x, y): pull from x - c_k with the wrap,
// bounce-back and stale rules above.  Stale entries are read from the node's own
// slots of D, which only this node's thread writes (callers gathering for another
// node never meet a stale entry; the host validates that).
template <typename T>
__device__ __forceinline__ void b_gather(const BatchLat &L, const T *__restrict__ S,
                                         const T *D, int x, int y, T f[Q])
{
    const int Nx = L.Nx, Ny = L.Ny;
    const unsigned P = L.plane;
    const unsigned nc = (unsigned)(y * Nx + x);
    const bool px = L.flags & BF_PER_X, py = L.flags & BF_PER_Y;
    f[0] = S[nc];
#pragma unroll
    for (int k = 1; k < Q; k++) {
        const int cx = D2Q9::cx[k], cy = D2Q9::cy[k];
        int xs = x - cx, ys = y - cy;
        bool oob = false;
        if (cx != 0) {
            if (xs < 0) {
                xs += Nx;
                oob = !px;
            } else if (xs >= Nx) {
                xs -= Nx;
                oob = !px;
            }
        }
        if (cy != 0) {
            if (ys < 0) {
                ys += Ny;
                oob = oob || !py;
            } else if (ys >= Ny) {
                ys -= Ny;
                oob = oob || !py;
            }
        }
        const bool bb = !oob && b_solid(L, xs, ys);
        const unsigned idx = oob ? k * P + nc
                                 : (bb ? D2Q9::opp[k] * P + nc : k * P + (unsigned)(ys * Nx + xs));
        f[k] = oob ? D[idx] : S[idx];
    }
}

__device__ __forceinline__ void b_store(double *p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void b_store(float *p, float v) { __builtin_nontemporal_store(v, p); }

// One node of one step s of lattice L.
template <typename T>
__device__ __forceinline__ void batch_node(const BatchLat &L, int s, int n)
{
    const int Nx = L.Nx, Ny = L.Ny;
    const int y = n / Nx, x = n - y * Nx;
    const unsigned F = L.flags;
    const bool last = (s + 1 == L.steps);
    if (b_solid(L, x, y)) {
        if (last && (F & BF_MOMENTS)) {
            L.out[0][n] = NAN;
            L.out[1][n] = NAN;
            L.out[2][n] = NAN;
        }
        return;
    }
    const T *S = (const T *)L.buf[s & 1];
    T *D = (T *)L.buf[(s + 1) & 1];
    const double rho0 = L.rho0;

    // convective_{inlet,outlet}: the node takes the post-streaming populations of
    // its inner neighbour (convective_inlet.c:20-27); the impedance wall node
    // starts from those of the node above it (regularized closure input).
    int gx = x, gy = y;
    const bool conv = (F & BF_CONVECTIVE) && (x == 0 || x == Nx - 1);
    if (conv) gx = (x == 0) ? 1 : Nx - 2;
    const bool liner = (F & BF_LINER) && y == 0 && x >= L.gap_x0 && x <= L.gap_x1;
    if (liner) gy = 1;
    T f[Q];
    b_gather(L, S, D, gx, gy, f);

    if (liner) {
        // cases/nit/nit.c:13-47 per node, then velocity_bottom_regularized.c:22-70
        const int e = x - L.gap_x0;
        double *st = L.state + 8 * e;
        TdibcState ts;
        ts.phi[0] = st[0];
        ts.phi[1] = st[1];
        ts.psi[0] = st[2];
        ts.psi[1] = st[3];
        ts.chi[0] = st[4];
        ts.chi[1] = st[5];
        const TdibcParams &P = L.tp;
        const double p = P.cs * P.cs * (st[6] - P.rho0) * P.cp;
        const double v = ade_velocity(ts, P, p);
        st[0] = ts.phi[0];
        st[1] = ts.phi[1];
        st[2] = ts.psi[0];
        st[3] = ts.psi[1];
        st[4] = ts.chi[0];
        st[5] = ts.chi[1];
        const double vbc = P.gain * (v / P.cu);
        const double *lg = L.saved[s & 1] + 3 * e;
        double fi[Q], fb[Q];
#pragma unroll
        for (int k = 0; k < Q; k++) fi[k] = b_dec(f[k], k, rho0);
        regularized_node(fi, lg[0], lg[1], lg[2], lg[0], 0.0, vbc, fb);
#pragma unroll
        for (int k = 0; k < Q; k++) f[k] = b_enc<T>(fb[k], k, rho0);
    }
    if ((F & BF_TOP_VELOCITY) && y == Ny - 1) {
        double fd[Q];
#pragma unroll
        for (int k = 0; k < Q; k++) fd[k] = b_dec(f[k], k, rho0);
        zou_he_top_node(fd, L.U_top, L.V_top);
        f[4] = b_enc<T>(fd[4], 4, rho0);
        f[7] = b_enc<T>(fd[7], 7, rho0);
        f[8] = b_enc<T>(fd[8], 8, rho0);
    }

    const bool force = F & BF_FORCE;
    const double om = L.om_col[x];
    double rl, ul, vl;   // next step's lagged moments as the reference's arrays hold them
    if constexpr (sizeof(T) == 8) {
        double *fd = reinterpret_cast<double *>(f);
        const Moments m = moments_ref(fd, L.Fx, L.Fy);
        rl = m.rho;
        ul = m.ux;
        vl = m.uy;
        if (last && (F & BF_MOMENTS)) {
            L.out[0][n] = m.rho;
            L.out[1][n] = m.ux;
            L.out[2][n] = m.uy;
        }
        collide_ref(fd, m, om, force, L.Fx, L.Fy);
    } else {
        float *q = reinterpret_cast<float *>(f);
        const float r0 = (float)rho0;
        const MomentsF m = moments_f32(q, r0, (float)L.Fx, (float)L.Fy);
        rl = (double)m.dr + rho0;
        ul = (double)m.ux;
        vl = (double)m.uy;
        if (last && (F & BF_MOMENTS)) {
            double fd[Q];
#pragma unroll
            for (int k = 0; k < Q; k++) fd[k] = b_dec(q[k], k, rho0);
            const Moments md = moments_ref(fd, L.Fx, L.Fy);
            L.out[0][n] = md.rho;
            L.out[1][n] = md.ux;
            L.out[2][n] = md.uy;
        }
        collide_f32(q, m, r0, (float)om, force, (float)L.Fx, (float)L.Fy);
    }

    if (F & BF_LINER) {
        if (liner) L.state[8 * (x - L.gap_x0) + 6] = rl;
        if (y == 1 && x >= L.gap_x0 && x <= L.gap_x1) {
            double *sv = L.saved[(s + 1) & 1] + 3 * (x - L.gap_x0);
            sv[0] = rl;
            sv[1] = ul;
            sv[2] = vl;
        }
    }
    for (int i = 0; i < L.nprobe; i++)
        if (n == L.probe_n[i]) L.probe_hist[(long long)(s + 1) * L.nprobe + i] = rl;

    const unsigned P = L.plane;
#pragma unroll
    for (int k = 0; k < Q; k++) b_store(&D[k * P + (unsigned)n], f[k]);
}

// One time step of every lattice whose own step lies in [0, steps).  Persistent:
// workgroup b owns tiles [t0, t1) of the tile table; with `banded` the blocks
// that the hardware deals to one XCD (b % 8) own one contiguous band of tiles, so
// a lattice's neighbouring rows meet in that XCD's L2.
template <typename T>
__global__ void __launch_bounds__(kBatchThreads) k_batch_step(const BatchLat *__restrict__ lats,
                                                             const Tile *__restrict__ tiles,
                                                             const BatchCtl *__restrict__ ctl,
                                                             int local, int tile, int banded)
{
    const int ntiles = ctl->ntiles;
    const long long epoch = ctl->epoch + local;
    const int G = gridDim.x, b = blockIdx.x;
    const int chunk = banded ? (b % 8) * (G / 8) + b / 8 : b;
    const int t0 = (int)((long long)ntiles * chunk / G);
    const int t1 = (int)((long long)ntiles * (chunk + 1) / G);
    for (int t = t0; t < t1; t++) {
        const Tile tl = tiles[t];
        const BatchLat &L = lats[tl.lat];
        const long long s = epoch - L.start;
        if (s < 0 || s >= L.steps) continue;
        const int N = L.Nx * L.Ny;
        const int end = min(tl.start + tile, N);
        for (int n = tl.start + (int)threadIdx.x; n < end; n += (int)blockDim.x)
            batch_node<T>(L, (int)s, n);
    }
}

static __global__ void k_batch_advance(BatchCtl *ctl, int n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) ctl->epoch += n;
}

// Initial state of one lattice: buf[1] = feq(M0) (what the reference's f_new holds
// before the first step), buf[0] = collide(feq(M0)) with the lagged moments M0,
// probe row 0, the impedance state and the first lagged row-1 moments.
template <typename T>
__global__ void __launch_bounds__(256) k_batch_init(const BatchLat *__restrict__ lats, int slot,
                                                    BatchInit in)
{
    const BatchLat &L = lats[slot];
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= L.Nx) return;
    const int n = y * L.Nx + x;
    double rho, ux, uy;
    if (in.kind == 0) {
        rho = in.prof[x];
        ux = 0.0;
        uy = 0.0;
    } else if (in.kind == 1) {
        // lbmx_init_taylor_green's formulas on host-computed sin/cos tables
        const double sx = in.tx[x], cxx = in.tx[L.Nx + x], c2x = in.tx[2 * L.Nx + x];
        const double sy = in.ty[y], cyy = in.ty[L.Ny + y], c2y = in.ty[2 * L.Ny + y];
        ux = in.u0 * sx * cyy;
        uy = -in.u0 * in.r * cxx * sy;
        rho = 1.0 - 0.75 * in.u0 * in.u0 * (c2x + in.r * in.r * c2y);
    } else {
        const size_t N = (size_t)L.Nx * L.Ny;
        rho = in.fields[n];
        ux = in.fields[N + n];
        uy = in.fields[2 * N + n];
    }
    const double rho0 = L.rho0;
    const double usq = ux * ux + uy * uy;
    T *A = (T *)L.buf[0], *B = (T *)L.buf[1];
    const unsigned P = L.plane;
    T q[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) {
        const double fe = feq_ref(D2Q9::w[k], (double)D2Q9::cx[k], (double)D2Q9::cy[k], rho, ux,
                                  uy, usq);
        q[k] = b_enc<T>(fe, k, rho0);
        B[k * P + n] = q[k];
    }
    const bool solid = b_solid(L, x, y);
    if (!solid) {
        const bool force = L.flags & BF_FORCE;
        const double om = L.om_col[x];
        if constexpr (sizeof(T) == 8) {
            collide_ref(reinterpret_cast<double *>(q), Moments{rho, ux, uy}, om, force, L.Fx,
                        L.Fy);
        } else {
            MomentsF m;
            m.dr = (float)(rho - rho0);
            m.ux = (float)ux;
            m.uy = (float)uy;
            collide_f32(reinterpret_cast<float *>(q), m, (float)rho0, (float)om, force,
                        (float)L.Fx, (float)L.Fy);
        }
    }
#pragma unroll
    for (int k = 0; k < Q; k++) A[k * P + n] = q[k];
    for (int i = 0; i < L.nprobe; i++)
        if (n == L.probe_n[i]) L.probe_hist[i] = rho;
    if ((L.flags & BF_LINER) && x >= L.gap_x0 && x <= L.gap_x1) {
        const int e = x - L.gap_x0;
        if (y == 0) {
            double *st = L.state + 8 * e;
            for (int i = 0; i < 6; i++) st[i] = 0.0;
            st[6] = rho;
            st[7] = 0.0;
        } else if (y == 1) {
            double *sv = L.saved[0] + 3 * e;
            sv[0] = rho;
            sv[1] = ux;
            sv[2] = uy;
        }
    }
}

// ---- spectra of a finished request's probes (server) ------------------------------
// Block (f, p): P(f) = sum_t p'(t) e^{-2 pi i f t} over the first `rows[p]`
// samples of probe `probe[p]`, p' = cs2 (rho - rho0).  Fixed-order reduction, so
// the result does not depend on anything but the history.
constexpr int kSpecThreads = 256;

struct SpecArgs {
    const double *hist;
    int nprobe;
    int probe[2];
    int rows[2];
    double rho0, cs2, f_lo, df;
    double *out;             // nfreq x 2 x (re, im)
};

static __global__ void __launch_bounds__(kSpecThreads) k_spectrum(SpecArgs a)
{
    __shared__ double lre[kSpecThreads], lim[kSpecThreads];
    const int fi = blockIdx.x, p = blockIdx.y;
    const double f = a.f_lo + a.df * fi;
    double re = 0.0, im = 0.0;
    for (int t = threadIdx.x; t < a.rows[p]; t += kSpecThreads) {
        const double v = a.cs2 * (a.hist[(long long)t * a.nprobe + a.probe[p]] - a.rho0);
        double sn, cs;
        sincospi(2.0 * f * (double)t, &sn, &cs);
        re += v * cs;
        im -= v * sn;
    }
    lre[threadIdx.x] = re;
    lim[threadIdx.x] = im;
    __syncthreads();
    for (int h = kSpecThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            lre[threadIdx.x] += lre[threadIdx.x + h];
            lim[threadIdx.x] += lim[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.out[4 * fi + 2 * p] = lre[0];
        a.out[4 * fi + 2 * p + 1] = lim[0];
    }
}

}  // namespace lbmx
