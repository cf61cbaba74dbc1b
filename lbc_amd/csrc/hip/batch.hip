// batch.hip -- host side of the batched engine (include/lbmx_batch.h): device
// arena, per-lattice descriptors, the tile table, HIP-graph chunks and the
// launch-configuration autotuner.
//
// Ordering model: everything runs on the batch's one stream.  Descriptor and
// tile-table updates are stream-ordered copies from pinned staging buffers, so a
// lattice that joins or leaves between chunks never races with kernels already
// queued; memory released by a finished lattice is reused only by work queued
// after it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "batch_impl.h"

using namespace lbmx;

// ---- arena ---------------------------------------------------------------------
int BatchArena::init(size_t bytes)
{
    cap = (bytes + 255) & ~(size_t)255;
    hipError_t e = hipMalloc((void **)&base, cap);
    if (e != hipSuccess) {
        base = nullptr;
        return lbmx_fail(LBMX_ERR_NOMEM, "batch arena of %zu bytes: %s", cap, hipGetErrorString(e));
    }
    free_.clear();
    free_[0] = cap;
    used = peak = 0;
    return LBMX_OK;
}

void *BatchArena::alloc(size_t bytes)
{
    bytes = (bytes + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        const size_t off = it->first, sz = it->second;
        free_.erase(it);
        if (sz > bytes) free_[off + bytes] = sz - bytes;
        used += bytes;
        peak = std::max(peak, used);
        return base + off;
    }
    return nullptr;
}

void BatchArena::release(void *p, size_t bytes)
{
    bytes = (bytes + 255) & ~(size_t)255;
    size_t off = (size_t)((char *)p - base);
    used -= bytes;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + bytes == next->first) {
        bytes += next->second;
        free_.erase(next);
    }
    auto it = free_.lower_bound(off);
    if (it != free_.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == off) {
            prev->second += bytes;
            return;
        }
    }
    free_[off] = bytes;
}

void BatchArena::destroy()
{
    if (base) hipFree(base);
    base = nullptr;
}

// ---- layout of one lattice's block ---------------------------------------------------
static size_t a256(size_t b) { return (b + 255) & ~(size_t)255; }

static BatchLayout layout_of(size_t esize, const lbmx_case *c)
{
    BatchLayout L{};
    const size_t N = (size_t)c->Nx * c->Ny;
    L.plane = (unsigned)(((N + 63) / 64) * 64 + 64);
    L.gap = (c->liner && c->gap_x1 >= c->gap_x0) ? c->gap_x1 - c->gap_x0 + 1 : 0;
    size_t o = 0;
    for (int i = 0; i < 2; i++) {
        L.buf[i] = o;
        o += a256((size_t)kQ * L.plane * esize);
    }
    // frame buffer of pair mode: the nodes within 3 of an edge
    L.fplane = (unsigned)(((frame_nodes_of(c->Nx, c->Ny) + 63) / 64) * 64 + 64);
    L.fbuf = o;
    o += a256((size_t)kQ * L.fplane * esize);
    L.om = o;
    o += a256((size_t)c->Nx * 8);
    for (int i = 0; i < 3; i++) {
        L.out[i] = o;
        if (c->keep_moments) o += a256(N * 8);
    }
    for (int i = 0; i < 2; i++) {
        L.saved[i] = o;
        o += a256((size_t)3 * L.gap * 8);
    }
    L.state = o;
    o += a256((size_t)8 * L.gap * 8);
    L.probe = o;
    o += a256((size_t)(c->steps + 1) * c->nprobe * 8);
    L.init_a = o;
    if (c->init == LBMX_INIT_PULSE) o += a256((size_t)c->Nx * 8);
    else if (c->init == LBMX_INIT_TGV) o += a256((size_t)3 * c->Nx * 8);
    else o += a256(3 * N * 8);
    L.init_b = o;
    if (c->init == LBMX_INIT_TGV) o += a256((size_t)3 * c->Ny * 8);
    L.total = o;
    return L;
}

// Whether the impedance ODE decays under the step of cases/nit/nit.c:19-35 (dt = 1):
// a real pole updates phi <- (1 + lambda) phi, stable for |1 + lambda| < 1; a
// complex pair updates psi and then chi with the new psi, amplification
// [[1+a, -b], [b(1+a), 1+a-b^2]], stable when det = (1+a)^2 < 1 and
// |trace| = |2(1+a) - b^2| < 1 + det (Jury).  Unused modes (zero weights) are free.
static bool ade_stable(const lbmx_impedance &z)
{
    const double tol = 1e-9;
    for (int i = 0; i < 2; i++) {
        if (z.A[i] != 0.0 && !(std::fabs(1.0 + z.lambdas[i]) < 1.0 - tol)) return false;
        if (z.B[i] != 0.0 || z.C[i] != 0.0) {
            const double a1 = 1.0 + z.alpha[i], b = z.beta[i];
            const double det = a1 * a1, tr = 2.0 * a1 - b * b;
            if (!(det < 1.0 - tol && std::fabs(tr) < 1.0 + det - tol)) return false;
        }
    }
    return true;
}

static int validate(const lbmx_case *c)
{
    if (c->Nx < 4 || c->Ny < 4) return lbmx_fail(LBMX_ERR_ARG, "case %dx%d: sizes must be >= 4", c->Nx, c->Ny);
    if ((long long)c->Nx * c->Ny > (1LL << 26))
        return lbmx_fail(LBMX_ERR_ARG, "case %dx%d: more than 2^26 nodes", c->Nx, c->Ny);
    if (!(c->tau > 0.5)) return lbmx_fail(LBMX_ERR_ARG, "tau must exceed 0.5");
    if (c->sponge < 0 || 2 * c->sponge > c->Nx || (c->sponge > 0 && !(c->tau_sponge > 0.5)))
        return lbmx_fail(LBMX_ERR_ARG, "invalid sponge");
    if (c->steps < 1) return lbmx_fail(LBMX_ERR_ARG, "steps must be >= 1");
    if (c->init < LBMX_INIT_PULSE || c->init > LBMX_INIT_FIELDS)
        return lbmx_fail(LBMX_ERR_ARG, "unknown init kind %d", c->init);
    if (c->init == LBMX_INIT_PULSE && c->amp != 0.0 && !(c->halfwidth > 0.0))
        return lbmx_fail(LBMX_ERR_ARG, "pulse half-width must be positive");
    if (c->walls & ~15u) return lbmx_fail(LBMX_ERR_ARG, "unknown wall bits");
    if (c->nprobe < 0 || c->nprobe > 4) return lbmx_fail(LBMX_ERR_ARG, "at most 4 probes");
    BatchLat d{};
    d.Nx = c->Nx;
    d.Ny = c->Ny;
    d.gap_x0 = (c->walls & LBMX_WALL_BOTTOM) ? c->gap_x0 : 0;
    d.gap_x1 = (c->walls & LBMX_WALL_BOTTOM) ? c->gap_x1 : -1;
    d.flags = ((c->walls & LBMX_WALL_BOTTOM) ? BF_WALL_BOTTOM : 0) |
              ((c->walls & LBMX_WALL_TOP) ? BF_WALL_TOP : 0) |
              ((c->walls & LBMX_WALL_LEFT) ? BF_WALL_LEFT : 0) |
              ((c->walls & LBMX_WALL_RIGHT) ? BF_WALL_RIGHT : 0);
    auto solid = [&](int x, int y) {
        const unsigned F = d.flags;
        return ((F & BF_WALL_BOTTOM) && y == 0 && (x < d.gap_x0 || x > d.gap_x1)) ||
               ((F & BF_WALL_TOP) && y == d.Ny - 1) || ((F & BF_WALL_LEFT) && x == 0) ||
               ((F & BF_WALL_RIGHT) && x == d.Nx - 1);
    };
    for (int i = 0; i < c->nprobe; i++) {
        if (c->probe_x[i] < 0 || c->probe_x[i] >= c->Nx || c->probe_y[i] < 0 ||
            c->probe_y[i] >= c->Ny)
            return lbmx_fail(LBMX_ERR_ARG, "probe %d outside the lattice", i);
        if (solid(c->probe_x[i], c->probe_y[i]))
            return lbmx_fail(LBMX_ERR_ARG, "probe %d on a solid node", i);
    }
    if (c->convective_ends) {
        // the copied column must be fluid wherever the end node is, and no end node
        // of a non-periodic row may be fluid (it would copy a stale entry)
        if (c->per_x) return lbmx_fail(LBMX_ERR_ARG, "convective ends need a non-periodic x axis");
        if (c->top_velocity)
            return lbmx_fail(LBMX_ERR_ARG, "convective ends with a moving top wall are not supported");
        if (!c->per_y)
            for (int y : {0, c->Ny - 1})
                for (int x : {0, c->Nx - 1})
                    if (!solid(x, y))
                        return lbmx_fail(LBMX_ERR_ARG,
                                         "convective ends: corner (%d,%d) must be solid", x, y);
        for (int y = 0; y < c->Ny; y++) {
            if (!solid(0, y) && solid(1, y))
                return lbmx_fail(LBMX_ERR_ARG, "convective inlet copies a solid node");
            if (!solid(c->Nx - 1, y) && solid(c->Nx - 2, y))
                return lbmx_fail(LBMX_ERR_ARG, "convective outlet copies a solid node");
        }
    }
    if (c->liner) {
        if (!(c->walls & LBMX_WALL_BOTTOM) || c->per_y)
            return lbmx_fail(LBMX_ERR_ARG, "the impedance wall is an opening of a bottom wall");
        // a non-periodic row keeps its end nodes solid (corners of the side walls or
        // of convective ends); a periodic row may be an impedance wall end to end
        const int lo = c->per_x ? 0 : 1, hi = c->per_x ? c->Nx - 1 : c->Nx - 2;
        if (c->gap_x1 < c->gap_x0 || c->gap_x0 < lo || c->gap_x1 > hi)
            return lbmx_fail(LBMX_ERR_ARG, "impedance opening must lie inside [%d, %d]", lo, hi);
        for (int x = c->gap_x0; x <= c->gap_x1; x++)
            if (solid(x, 1)) return lbmx_fail(LBMX_ERR_ARG, "impedance opening under a solid node");
        if (!ade_stable(c->imp))
            return lbmx_fail(LBMX_ERR_ARG, "impedance ODE grows under the explicit step "
                                           "(a real pole outside (-2, 0) or an unstable pair)");
    } else if ((c->walls & LBMX_WALL_BOTTOM) && c->gap_x1 >= c->gap_x0) {
        return lbmx_fail(LBMX_ERR_ARG, "a bottom-wall opening needs the impedance wall");
    }
    if (c->top_velocity && (c->per_y || (c->walls & LBMX_WALL_TOP)))
        return lbmx_fail(LBMX_ERR_ARG, "velocity_top needs a non-periodic, open top row");
    return LBMX_OK;
}

// ---- staging uploads ----------------------------------------------------------------
static int upload(lbmx_batch *b, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return LBMX_OK;
    BatchStaging &s = b->ring[b->ring_next];
    b->ring_next = (b->ring_next + 1) % (int)b->ring.size();
    if (s.pending) {
        HIPCHK(hipEventSynchronize(s.ev));
        s.pending = false;
    }
    if (s.cap < bytes) {
        if (s.p) HIPCHK(hipHostFree(s.p));
        s.p = nullptr;
        const size_t cap = std::max(bytes, (size_t)1 << 16);
        HIPCHK(hipHostMalloc(&s.p, cap, hipHostMallocDefault));
        s.cap = cap;
    }
    std::memcpy(s.p, src, bytes);
    HIPCHK(hipMemcpyAsync(dst, s.p, bytes, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipEventRecord(s.ev, b->stream));
    s.pending = true;
    return LBMX_OK;
}

// ---- descriptors -----------------------------------------------------------------------
static void fill_desc(lbmx_batch *b, BatchSlot &S)
{
    const lbmx_case *c = &S.c;
    const BatchLayout &L = S.lay;
    char *blk = S.block;
    BatchLat &d = S.d;
    d = BatchLat{};
    d.buf[0] = blk + L.buf[0];
    d.buf[1] = blk + L.buf[1];
    d.fbuf = blk + L.fbuf;
    d.fplane = L.fplane;
    d.om_col = (const double *)(blk + L.om);
    for (int i = 0; i < 3; i++) d.out[i] = c->keep_moments ? (double *)(blk + L.out[i]) : nullptr;
    d.saved[0] = (double *)(blk + L.saved[0]);
    d.saved[1] = (double *)(blk + L.saved[1]);
    d.state = (double *)(blk + L.state);
    d.probe_hist = (double *)(blk + L.probe);
    d.status = b->status_dev + (&S - b->slots.data());
    d.start = -1;
    d.plane = L.plane;
    d.Nx = c->Nx;
    d.Ny = c->Ny;
    unsigned F = 0;
    if (c->per_x) F |= BF_PER_X;
    if (c->per_y) F |= BF_PER_Y;
    if (c->walls & LBMX_WALL_BOTTOM) F |= BF_WALL_BOTTOM;
    if (c->walls & LBMX_WALL_TOP) F |= BF_WALL_TOP;
    if (c->walls & LBMX_WALL_LEFT) F |= BF_WALL_LEFT;
    if (c->walls & LBMX_WALL_RIGHT) F |= BF_WALL_RIGHT;
    if (c->top_velocity) F |= BF_TOP_VELOCITY;
    if (c->convective_ends) F |= BF_CONVECTIVE;
    if (c->liner) F |= BF_LINER;
    if (c->Fx != 0.0 || c->Fy != 0.0) F |= BF_FORCE;
    if (c->keep_moments) F |= BF_MOMENTS;
    d.flags = F;
    d.gap_x0 = (c->walls & LBMX_WALL_BOTTOM) ? c->gap_x0 : 0;
    d.gap_x1 = (c->walls & LBMX_WALL_BOTTOM) ? c->gap_x1 : -1;
    d.steps = c->steps;
    d.nprobe = c->nprobe;
    for (int i = 0; i < c->nprobe; i++) d.probe_n[i] = c->probe_y[i] * c->Nx + c->probe_x[i];
    d.inv_nx = 1.0 / c->Nx;
    d.rho0 = c->rho0;
    d.Fx = c->Fx;
    d.Fy = c->Fy;
    d.U_top = c->U_top;
    d.V_top = c->V_top;
    TdibcParams &P = d.tp;
    P.rho0 = c->rho0;
    P.cs = 1.0 / std::sqrt(3.0);
    P.cp = 1.0;
    P.dt = 1.0;
    P.cu = 1.0;
    P.gain = c->gain;
    P.Zinf = c->imp.Zinf;
    for (int i = 0; i < 2; i++) {
        P.A[i] = c->imp.A[i];
        P.lambdas[i] = c->imp.lambdas[i];
        P.B[i] = c->imp.B[i];
        P.C[i] = c->imp.C[i];
        P.alpha[i] = c->imp.alpha[i];
        P.beta[i] = c->imp.beta[i];
    }
    S.init = BatchInit{};
    S.init.kind = c->init;
    S.init.prof = (const double *)(blk + L.init_a);
    S.init.tx = (const double *)(blk + L.init_a);
    S.init.ty = (const double *)(blk + L.init_b);
    S.init.fields = (const double *)(blk + L.init_a);
    S.init.u0 = c->u0;
    const double kx = 2.0 * M_PI / c->Nx, ky = 2.0 * M_PI / c->Ny;
    S.init.r = kx / ky;
}

// Host-side set-up tables (libm, same formulas as the reference drivers' set-up).
static int upload_tables(lbmx_batch *b, BatchSlot &S)
{
    const lbmx_case *c = &S.c;
    std::vector<double> om(c->Nx);
    for (int x = 0; x < c->Nx; x++) {
        double tau = c->tau;
        const int dd = std::min(x, c->Nx - 1 - x);   // distance to the nearer end
        if (c->sponge > 0 && dd < c->sponge)
            tau = c->tau_sponge + (c->tau - c->tau_sponge) * ((double)dd / c->sponge);
        om[x] = 1.0 / tau;
    }
    if (int rc = upload(b, S.block + S.lay.om, om.data(), om.size() * 8)) return rc;
    if (c->init == LBMX_INIT_PULSE) {
        std::vector<double> prof(c->Nx);
        for (int x = 0; x < c->Nx; x++) {
            if (c->amp == 0.0) {
                prof[x] = c->rho0;
            } else {
                const double xr = (x - c->x0) / c->halfwidth;
                prof[x] = c->rho0 + c->amp * std::exp(-std::log(2.0) * xr * xr);
            }
        }
        return upload(b, S.block + S.lay.init_a, prof.data(), prof.size() * 8);
    }
    if (c->init == LBMX_INIT_TGV) {
        const double kx = 2.0 * M_PI / c->Nx, ky = 2.0 * M_PI / c->Ny;
        std::vector<double> tx(3 * c->Nx), ty(3 * c->Ny);
        for (int x = 0; x < c->Nx; x++) {
            tx[x] = std::sin(kx * x);
            tx[c->Nx + x] = std::cos(kx * x);
            tx[2 * c->Nx + x] = std::cos(2.0 * kx * x);
        }
        for (int y = 0; y < c->Ny; y++) {
            ty[y] = std::sin(ky * y);
            ty[c->Ny + y] = std::cos(ky * y);
            ty[2 * c->Ny + y] = std::cos(2.0 * ky * y);
        }
        if (int rc = upload(b, S.block + S.lay.init_a, tx.data(), tx.size() * 8)) return rc;
        return upload(b, S.block + S.lay.init_b, ty.data(), ty.size() * 8);
    }
    return LBMX_OK;   // LBMX_INIT_FIELDS: lbmx_batch_set_fields
}

// ---- launches ---------------------------------------------------------------------------
static int grid_of(const lbmx_batch *b)
{
    if (b->tun.order == 5) return std::max(8, b->grid5);
    int G = b->ncu * b->tun.groups_per_cu;
    return std::max(8, (G / 8) * 8);
}

// Interior grid of tile order 5 for n tiles: n rounded up to a multiple of the
// largest power of two <= n / 8 (at most 1/8 idle workgroups), so the grid -- and
// with it the captured graphs -- changes only when n crosses such a step.
static int grid5_of(int n)
{
    int q = 8;
    while (q * 16 <= n) q *= 2;
    return std::max(q, (n + q - 1) / q * q);
}

template <typename T, bool W>
static void launch_step_t(lbmx_batch *b, int local)
{
    const dim3 grid(grid_of(b)), block(b->tun.threads);
    const dim3 fgrid(2 * b->ncu), fblock(kFrameThreads);
    hipLaunchKernelGGL((k_batch_interior<T, W>), grid, block, 0, b->stream, b->lats_dev,
                       b->tiles_dev, b->ctl_dev, local, b->tun.tile, b->tun.order);
    hipLaunchKernelGGL((k_batch_frame<T, W>), fgrid, fblock, 0, b->stream, b->lats_dev,
                       b->tiles_dev, b->ctl_dev, local);
}

// One pair of steps (pair mode): band step 1, core steps 1-2, band step 2.
template <typename T, bool W>
static void launch_pair_t(lbmx_batch *b, int local)
{
    const dim3 gc(std::max(8, b->grid5)), gb(std::max(8, b->grid_band));
    hipLaunchKernelGGL((k_pair_band<T, W, 0>), gb, dim3(kBandTile), 0, b->stream, b->lats_dev,
                       b->tiles_dev, b->ctl_dev, local);
    hipLaunchKernelGGL((k_pair_core<T, W>), gc, dim3(PairShape<T>::threads), 0, b->stream,
                       b->lats_dev, b->tiles_dev, b->ctl_dev, local);
    hipLaunchKernelGGL((k_pair_band<T, W, 1>), gb, dim3(kBandTile), 0, b->stream, b->lats_dev,
                       b->tiles_dev, b->ctl_dev, local);
}

static int launch_step(lbmx_batch *b, int local)
{
    if (b->pairs) {
        // local is the first step of a pair (even)
        if (b->precision == LBMX_F64)
            launch_pair_t<double, false>(b, local);
        else if (b->precision == LBMX_F32_WIDE)
            launch_pair_t<float, true>(b, local);
        else
            launch_pair_t<float, false>(b, local);
        HIPCHK(hipGetLastError());
        return LBMX_OK;
    }
    if (b->precision == LBMX_F64)
        launch_step_t<double, false>(b, local);
    else if (b->precision == LBMX_F32_WIDE)
        launch_step_t<float, true>(b, local);
    else
        launch_step_t<float, false>(b, local);
    HIPCHK(hipGetLastError());
    return LBMX_OK;
}

static int launch_advance(lbmx_batch *b, int n)
{
    hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(64), 0, b->stream, b->ctl_dev, n);
    HIPCHK(hipGetLastError());
    return LBMX_OK;
}

static int launch_init(lbmx_batch *b, int slot)
{
    BatchSlot &S = b->slots[slot];
    const dim3 grid((S.c.Nx + 255) / 256, S.c.Ny), block(256);
    if (b->precision == LBMX_F64)
        hipLaunchKernelGGL((k_batch_init<double, false>), grid, block, 0, b->stream, b->lats_dev,
                           slot, S.init);
    else if (b->precision == LBMX_F32_WIDE)
        hipLaunchKernelGGL((k_batch_init<float, true>), grid, block, 0, b->stream, b->lats_dev,
                           slot, S.init);
    else
        hipLaunchKernelGGL((k_batch_init<float, false>), grid, block, 0, b->stream, b->lats_dev,
                           slot, S.init);
    HIPCHK(hipGetLastError());
    return LBMX_OK;
}

// Destroys the retired graphs; the stream must be idle.
static void reap_graphs(lbmx_batch *b)
{
    for (hipGraphExec_t g : b->retired) hipGraphExecDestroy(g);
    b->retired.clear();
}

// Retires the captured graphs (a launch of them may still be queued); they are
// destroyed at the next point where the stream is idle.
static void drop_graphs(lbmx_batch *b)
{
    for (auto &kv : b->graphs) b->retired.push_back(kv.second);
    b->graphs.clear();
    if (b->retired.size() > 32 && hipStreamSynchronize(b->stream) == hipSuccess) reap_graphs(b);
}

// Starts waiting lattices at the current epoch and rewrites the tile table.
static int commit(lbmx_batch *b)
{
    if (!b->dirty) return LBMX_OK;
    for (int slot : b->order) {
        BatchSlot &S = b->slots[slot];
        if (!S.pending) continue;
        if (S.c.init == LBMX_INIT_FIELDS && !S.fields_set)
            return lbmx_fail(LBMX_ERR_STATE, "lattice %d: initial fields not set", slot);
        S.d.start = b->epoch;
        if (int rc = upload(b, &b->lats_dev[slot], &S.d, sizeof(BatchLat))) return rc;
        if (int rc = launch_init(b, slot)) return rc;
        S.pending = 0;
    }
    std::vector<Tile> &T = b->tile_host;
    T.clear();
    const int tile = b->tun.tile;
    if (b->pairs) {
        // core tiles (kPairTX x pair_ty nodes), then band tiles (kBandTile nodes)
        for (int slot : b->order) {
            const BatchSlot &S = b->slots[slot];
            const int nt = core_tiles(S.c.Nx, S.c.Ny, pair_ty((int)b->esize),
                                      wide_core(S.d.flags, S.c.Nx, S.c.Ny));
            for (int t = 0; t < nt; t++) T.push_back(Tile{slot, t});
        }
    } else {
        for (int slot : b->order) {
            const BatchSlot &S = b->slots[slot];
            const int ni = interior_nodes(S.c.Nx, S.c.Ny);
            for (int s0 = 0; s0 < ni; s0 += tile) T.push_back(Tile{slot, s0});
        }
    }
    const int n_interior = (int)T.size();
    for (int slot : b->order) {
        const BatchSlot &S = b->slots[slot];
        if (b->pairs) {
            const int nb = band_nodes(S.c.Nx, S.c.Ny, wide_core(S.d.flags, S.c.Nx, S.c.Ny));
            for (int b0 = 0; b0 < nb; b0 += kBandTile) T.push_back(Tile{slot, b0});
        } else {
            const int nf = frame_nodes(S.c.Nx, S.c.Ny);
            for (int f0 = 0; f0 < nf; f0 += kFrameTile) T.push_back(Tile{slot, f0});
        }
    }
    if ((long long)T.size() > b->tiles_cap)
        return lbmx_fail(LBMX_ERR_NOMEM, "tile table overflow (%zu > %lld)", T.size(), b->tiles_cap);
    if (int rc = upload(b, b->tiles_dev, T.data(), T.size() * sizeof(Tile))) return rc;
    BatchCtl ctl{b->epoch, n_interior, (int)T.size() - n_interior};
    if (int rc = upload(b, b->ctl_dev, &ctl, sizeof ctl)) return rc;
    const int g5 = grid5_of(n_interior);
    const int gb = grid5_of((int)T.size() - n_interior);
    if ((b->tun.order == 5 || b->pairs) && g5 != b->grid5) drop_graphs(b);
    if (b->pairs && gb != b->grid_band) drop_graphs(b);
    b->n_interior = n_interior;
    b->grid5 = g5;
    b->grid_band = gb;
    b->dirty = false;
    return LBMX_OK;
}

static int graph_for(lbmx_batch *b, int K, hipGraphExec_t *out)
{
    auto it = b->graphs.find(K);
    if (it != b->graphs.end()) {
        *out = it->second;
        return LBMX_OK;
    }
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = LBMX_OK;
    const int per = b->pairs ? 2 : 1;
    for (int l = 0; l < K && rc == LBMX_OK; l += per) rc = launch_step(b, l);
    if (rc == LBMX_OK) rc = launch_advance(b, K);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return lbmx_fail(LBMX_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return lbmx_fail(LBMX_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    b->graphs[K] = ex;
    *out = ex;
    return LBMX_OK;
}

int lbmx_batch_enqueue(lbmx_batch *b, int nsteps)
{
    if (b->pairs && (nsteps & 1))
        return lbmx_fail(LBMX_ERR_ARG, "pair mode advances in steps of two (asked for %d)", nsteps);
    if (int rc = commit(b)) return rc;
    const int K = b->chunk;
    const int full = (b->use_graphs && K > 1) ? nsteps / K : 0;
    if (full > 0) {
        hipGraphExec_t ex;
        if (int rc = graph_for(b, K, &ex)) return rc;
        for (int i = 0; i < full; i++) HIPCHK(hipGraphLaunch(ex, b->stream));
    }
    const int rem = nsteps - full * K;
    for (int l = 0; l < rem; l += b->pairs ? 2 : 1)
        if (int rc = launch_step(b, l)) return rc;
    if (rem > 0)
        if (int rc = launch_advance(b, rem)) return rc;
    b->epoch += nsteps;
    return LBMX_OK;
}

const BatchSlot *lbmx_batch_slot(const lbmx_batch *b, int id)
{
    if (!b || id < 0 || id >= (int)b->slots.size() || !b->slots[id].used) return nullptr;
    return &b->slots[id];
}

int lbmx_batch_totals_of(const lbmx_batch *b, int id, double *scratch, double *out,
                         hipStream_t stream)
{
    const BatchSlot *S = lbmx_batch_slot(b, id);
    if (!S) return lbmx_fail(LBMX_ERR_ARG, "no lattice %d", id);
    if (!S->c.keep_moments) return lbmx_fail(LBMX_ERR_STATE, "lattice %d keeps no moments", id);
    int done = 0;
    lbmx_batch_steps_done(const_cast<lbmx_batch *>(b), id, &done);
    if (done < S->c.steps) return lbmx_fail(LBMX_ERR_STATE, "lattice %d has not completed", id);
    TotalsArgs a{};
    a.R = S->d.out[0];
    a.U = S->d.out[1];
    a.V = S->d.out[2];
    a.Nx = S->c.Nx;
    a.Ny = S->c.Ny;
    a.pitch = S->c.Nx;
    a.lat = b->lats_dev + id;
    HIPCHK(launch_totals(a, scratch, out, stream));
    return LBMX_OK;
}

int lbmx_tl_check(const lbmx_case *c, const lbmx_tl_spec *tl)
{
    if (tl->nfreq < 1 || tl->nfreq > 64) return lbmx_fail(LBMX_ERR_ARG, "nfreq must be in [1, 64]");
    if (tl->up_probe < 0 || tl->up_probe >= c->nprobe || tl->down_probe < 0 ||
        tl->down_probe >= c->nprobe)
        return lbmx_fail(LBMX_ERR_ARG, "spectrum probes must name probes of the case");
    if (tl->up_t0 < 0 || tl->up_t1 <= tl->up_t0 || tl->up_t1 > c->steps + 1 ||
        tl->down_t0 < 0 || tl->down_t1 <= tl->down_t0 || tl->down_t1 > c->steps + 1)
        return lbmx_fail(LBMX_ERR_ARG, "probe windows must lie inside [0, steps]");
    return LBMX_OK;
}

int lbmx_batch_spectrum_of(const BatchSlot &S, const lbmx_tl_spec &tl, hipStream_t stream,
                           double *dev, double *host, lbmx_result *out)
{
    SpecArgs a{};
    a.hist = S.d.probe_hist;
    a.nprobe = S.c.nprobe;
    a.probe[0] = tl.up_probe;
    a.probe[1] = tl.down_probe;
    a.t0[0] = tl.up_t0;
    a.t1[0] = tl.up_t1;
    a.t0[1] = tl.down_t0;
    a.t1[1] = tl.down_t1;
    a.rho0 = S.c.rho0;
    const double cs = 1.0 / std::sqrt(3.0);
    a.cs2 = cs * cs;
    a.f_lo = tl.f_lo;
    a.df = tl.nfreq > 1 ? (tl.f_hi - tl.f_lo) / (tl.nfreq - 1) : 0.0;
    a.out = dev;
    hipLaunchKernelGGL(k_spectrum, dim3(tl.nfreq, 2), dim3(kSpecThreads), 0, stream, a);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, dev, (size_t)tl.nfreq * 4 * 8, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    out->nfreq = tl.nfreq;
    for (int f = 0; f < tl.nfreq; f++) {
        const double *v = host + 4 * f;
        out->up_abs[f] = std::hypot(v[0], v[1]);
        out->down_abs[f] = std::hypot(v[2], v[3]);
        out->tl_db[f] = 20.0 * std::log10(out->up_abs[f] / out->down_abs[f]);
    }
    return LBMX_OK;
}

// ---- C ABI --------------------------------------------------------------------------
extern "C" {

LBMX_API size_t lbmx_case_bytes(int precision, const lbmx_case *c)
{
    if (!c) return 0;
    return layout_of(precision == LBMX_F64 ? 8 : 4, c).total;
}

LBMX_API int lbmx_case_check(const lbmx_case *c)
{
    if (!c) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    return validate(c);
}

LBMX_API int lbmx_batch_create(int precision, size_t budget, lbmx_batch **out)
{
    if (!out) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    *out = nullptr;
    if (precision != LBMX_F64 && precision != LBMX_F32_SHIFTED && precision != LBMX_F32_WIDE)
        return lbmx_fail(LBMX_ERR_ARG, "unknown precision %d", precision);
    if (budget < ((size_t)1 << 20)) return lbmx_fail(LBMX_ERR_ARG, "budget below 1 MiB");
    lbmx_batch *b = new lbmx_batch();
    b->precision = precision;
    b->esize = precision == LBMX_F64 ? 8 : 4;
    int rc = lbmx_device_info(&b->ncu);
    if (rc) {
        delete b;
        return rc;
    }
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return lbmx_fail(LBMX_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    // every tile but the last interior and the last frame tile of a lattice holds
    // >= 256 nodes of >= 18*esize bytes
    b->tiles_cap = (long long)(budget / (18 * b->esize * 256)) + 2 * kBatchMaxSlots + 1024;
    rc = b->arena.init(budget);
    if (!rc && hipMalloc((void **)&b->lats_dev, sizeof(BatchLat) * kBatchMaxSlots) != hipSuccess)
        rc = lbmx_fail(LBMX_ERR_NOMEM, "descriptor table");
    if (!rc && hipMalloc((void **)&b->tiles_dev, sizeof(Tile) * b->tiles_cap) != hipSuccess)
        rc = lbmx_fail(LBMX_ERR_NOMEM, "tile table");
    if (!rc && hipMalloc((void **)&b->ctl_dev, sizeof(BatchCtl)) != hipSuccess)
        rc = lbmx_fail(LBMX_ERR_NOMEM, "control block");
    if (!rc && hipMalloc((void **)&b->status_dev, sizeof(unsigned) * kBatchMaxSlots) != hipSuccess)
        rc = lbmx_fail(LBMX_ERR_NOMEM, "status words");
    if (!rc && hipMemset(b->status_dev, 0, sizeof(unsigned) * kBatchMaxSlots) != hipSuccess)
        rc = lbmx_fail(LBMX_ERR_HIP, "status words");
    if (!rc && hipMalloc((void **)&b->tot_dev, sizeof(double) * (2 * kTotMaxBlocks + 2)) !=
                   hipSuccess)
        rc = lbmx_fail(LBMX_ERR_NOMEM, "totals scratch");
    if (!rc) {
        b->ring.resize(64);   // uploads of a burst of admissions rarely wait on a busy slot
        for (auto &s : b->ring)
            if (hipEventCreateWithFlags(&s.ev, hipEventDisableTiming) != hipSuccess)
                rc = lbmx_fail(LBMX_ERR_HIP, "event");
    }
    if (!rc && (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess))
        rc = lbmx_fail(LBMX_ERR_HIP, "event");
    if (rc) {
        lbmx_batch_destroy(b);
        return rc;
    }
    BatchCtl ctl{0, 0, 0};
    if (hipMemcpy(b->ctl_dev, &ctl, sizeof ctl, hipMemcpyHostToDevice) != hipSuccess) {
        lbmx_batch_destroy(b);
        return lbmx_fail(LBMX_ERR_HIP, "control block upload");
    }
    *out = b;
    return LBMX_OK;
}

LBMX_API void lbmx_batch_destroy(lbmx_batch *b)
{
    if (!b) return;
    if (b->stream) hipStreamSynchronize(b->stream);
    for (auto &kv : b->graphs) b->retired.push_back(kv.second);
    b->graphs.clear();
    reap_graphs(b);
    for (auto &s : b->ring) {
        if (s.p) hipHostFree(s.p);
        if (s.ev) hipEventDestroy(s.ev);
    }
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->status_dev) hipFree(b->status_dev);
    if (b->tot_dev) hipFree(b->tot_dev);
    if (b->spec_dev) hipFree(b->spec_dev);
    if (b->spec_host) hipHostFree(b->spec_host);
    if (b->lats_dev) hipFree(b->lats_dev);
    if (b->tiles_dev) hipFree(b->tiles_dev);
    if (b->ctl_dev) hipFree(b->ctl_dev);
    b->arena.destroy();
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

LBMX_API int lbmx_batch_add(lbmx_batch *b, const lbmx_case *c, int *id)
{
    if (!b || !c || !id) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (int rc = validate(c)) return rc;
    int slot = -1;
    for (int i = 0; i < (int)b->slots.size(); i++)
        if (!b->slots[i].used) {
            slot = i;
            break;
        }
    if (slot < 0) {
        if ((int)b->slots.size() >= kBatchMaxSlots)
            return lbmx_fail(LBMX_ERR_NOMEM, "more than %d lattices", kBatchMaxSlots);
        b->slots.emplace_back();
        slot = (int)b->slots.size() - 1;
    }
    BatchSlot &S = b->slots[slot];
    S = BatchSlot{};
    S.c = *c;
    S.lay = layout_of(b->esize, c);
    S.block = (char *)b->arena.alloc(S.lay.total);
    if (!S.block)
        return lbmx_fail(LBMX_ERR_NOMEM, "budget: %zu bytes needed, %zu of %zu in use",
                         S.lay.total, b->arena.used, b->arena.cap);
    S.used = 1;
    S.pending = 1;
    fill_desc(b, S);
    if (int rc = upload_tables(b, S)) {
        b->arena.release(S.block, S.lay.total);
        S.used = 0;
        return rc;
    }
    b->order.push_back(slot);
    b->dirty = true;
    *id = slot;
    return LBMX_OK;
}

LBMX_API int lbmx_batch_set_fields(lbmx_batch *b, int id, const double *rho, const double *u,
                                   const double *v)
{
    BatchSlot *S = const_cast<BatchSlot *>(lbmx_batch_slot(b, id));
    if (!S || !rho || !u || !v) return lbmx_fail(LBMX_ERR_ARG, "invalid lattice or null array");
    if (S->c.init != LBMX_INIT_FIELDS || !S->pending)
        return lbmx_fail(LBMX_ERR_STATE, "fields can only be set before the first step of an "
                                         "LBMX_INIT_FIELDS lattice");
    const size_t N = (size_t)S->c.Nx * S->c.Ny;
    char *dst = S->block + S->lay.init_a;
    if (int rc = upload(b, dst, rho, N * 8)) return rc;
    if (int rc = upload(b, dst + N * 8, u, N * 8)) return rc;
    if (int rc = upload(b, dst + 2 * N * 8, v, N * 8)) return rc;
    S->fields_set = 1;
    return LBMX_OK;
}

LBMX_API int lbmx_batch_remove(lbmx_batch *b, int id)
{
    BatchSlot *S = const_cast<BatchSlot *>(lbmx_batch_slot(b, id));
    if (!S) return lbmx_fail(LBMX_ERR_ARG, "no lattice %d", id);
    b->arena.release(S->block, S->lay.total);
    S->used = 0;
    S->block = nullptr;
    b->order.erase(std::remove(b->order.begin(), b->order.end(), id), b->order.end());
    b->dirty = true;
    return LBMX_OK;
}

LBMX_API int lbmx_batch_run(lbmx_batch *b, int nsteps)
{
    if (!b || nsteps < 0) return lbmx_fail(LBMX_ERR_ARG, "invalid argument");
    return lbmx_batch_enqueue(b, nsteps);
}

LBMX_API int lbmx_batch_time(lbmx_batch *b, int nsteps, double *ms)
{
    if (!b || !ms || nsteps < 0) return lbmx_fail(LBMX_ERR_ARG, "invalid argument");
    if (int rc = commit(b)) return rc;
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    if (int rc = lbmx_batch_enqueue(b, nsteps)) return rc;
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    HIPCHK(hipEventSynchronize(b->ev1));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, b->ev0, b->ev1));
    *ms = t;
    return LBMX_OK;
}

LBMX_API int lbmx_batch_synchronize(lbmx_batch *b)
{
    if (!b) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(b->stream));
    reap_graphs(b);
    return LBMX_OK;
}

LBMX_API int lbmx_batch_steps_done(lbmx_batch *b, int id, int *steps)
{
    const BatchSlot *S = lbmx_batch_slot(b, id);
    if (!S || !steps) return lbmx_fail(LBMX_ERR_ARG, "no lattice %d", id);
    if (S->pending) {
        *steps = 0;
        return LBMX_OK;
    }
    const long long d = b->epoch - S->d.start;
    *steps = (int)std::min<long long>(std::max<long long>(d, 0), S->c.steps);
    return LBMX_OK;
}

LBMX_API int lbmx_batch_moments(lbmx_batch *b, int id, double *rho, double *u, double *v)
{
    const BatchSlot *S = lbmx_batch_slot(b, id);
    if (!S) return lbmx_fail(LBMX_ERR_ARG, "no lattice %d", id);
    if (!S->c.keep_moments) return lbmx_fail(LBMX_ERR_STATE, "lattice %d keeps no moments", id);
    int done = 0;
    lbmx_batch_steps_done(b, id, &done);
    if (done < S->c.steps) return lbmx_fail(LBMX_ERR_STATE, "lattice %d has not completed", id);
    HIPCHK(hipStreamSynchronize(b->stream));
    const size_t N = (size_t)S->c.Nx * S->c.Ny;
    double *dst[3] = {rho, u, v};
    for (int i = 0; i < 3; i++)
        if (dst[i]) HIPCHK(hipMemcpy(dst[i], S->d.out[i], N * 8, hipMemcpyDeviceToHost));
    return LBMX_OK;
}

LBMX_API int lbmx_batch_probes(lbmx_batch *b, int id, double *hist, int capacity)
{
    const BatchSlot *S = lbmx_batch_slot(b, id);
    if (!S || !hist || capacity < 0) return lbmx_fail(LBMX_ERR_ARG, "invalid argument");
    if (S->c.nprobe == 0 || S->pending) return 0;
    int done = 0;
    lbmx_batch_steps_done(b, id, &done);
    const int rows = std::min(capacity, done + 1);
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(hist, S->d.probe_hist, (size_t)rows * S->c.nprobe * 8,
                     hipMemcpyDeviceToHost));
    return rows;
}

LBMX_API int lbmx_batch_status(lbmx_batch *b, int id, unsigned *guard)
{
    const BatchSlot *S = lbmx_batch_slot(b, id);
    if (!S || !guard) return lbmx_fail(LBMX_ERR_ARG, "invalid argument");
    *guard = 0;
    if (S->pending) return LBMX_OK;
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(guard, b->status_dev + id, sizeof(unsigned), hipMemcpyDeviceToHost));
    return LBMX_OK;
}

LBMX_API int lbmx_batch_totals(lbmx_batch *b, int id, double *mass, double *energy)
{
    const BatchSlot *S = lbmx_batch_slot(b, id);
    if (!S || !mass || !energy) return lbmx_fail(LBMX_ERR_ARG, "invalid argument");
    if (int rc = lbmx_batch_totals_of(b, id, b->tot_dev, b->tot_dev + 2 * kTotMaxBlocks,
                                      b->stream))
        return rc;
    double t[2];
    HIPCHK(hipMemcpyAsync(t, b->tot_dev + 2 * kTotMaxBlocks, sizeof t, hipMemcpyDeviceToHost,
                          b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    *mass = t[0];
    *energy = t[1];
    return LBMX_OK;
}

LBMX_API int lbmx_batch_spectrum(lbmx_batch *b, int id, const lbmx_tl_spec *tl,
                                 lbmx_result *out)
{
    const BatchSlot *S = lbmx_batch_slot(b, id);
    if (!S || !tl || !out) return lbmx_fail(LBMX_ERR_ARG, "invalid argument");
    if (int rc = lbmx_tl_check(&S->c, tl)) return rc;
    int done = 0;
    lbmx_batch_steps_done(b, id, &done);
    if (done < S->c.steps) return lbmx_fail(LBMX_ERR_STATE, "lattice %d has not completed", id);
    if (!b->spec_dev) {
        HIPCHK(hipMalloc((void **)&b->spec_dev, 64 * 4 * 8));
        HIPCHK(hipHostMalloc((void **)&b->spec_host, 64 * 4 * 8, hipHostMallocDefault));
    }
    *out = lbmx_result{};
    out->tag = -1;
    out->steps = S->c.steps;
    return lbmx_batch_spectrum_of(*S, *tl, b->stream, b->spec_dev, b->spec_host, out);
}

LBMX_API int lbmx_batch_set_tuning(lbmx_batch *b, const lbmx_tuning *t)
{
    if (!b || !t) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if ((t->threads != 64 && t->threads != 128 && t->threads != 256) || t->tile < 256 ||
        t->tile % t->threads != 0 || t->tile / t->threads > 16 || t->groups_per_cu < 1 ||
        t->groups_per_cu > 32)
        return lbmx_fail(LBMX_ERR_ARG, "invalid tuning (threads %d, tile %d, groups/CU %d)",
                         t->threads, t->tile, t->groups_per_cu);
    if (t->order < 0 || t->order > 5) return lbmx_fail(LBMX_ERR_ARG, "tile order must be 0 .. 5");
    b->tun = *t;
    drop_graphs(b);
    b->dirty = true;
    return LBMX_OK;
}

LBMX_API int lbmx_batch_get_tuning(lbmx_batch *b, lbmx_tuning *t)
{
    if (!b || !t) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    *t = b->tun;
    return LBMX_OK;
}

LBMX_API int lbmx_batch_set_pairs(lbmx_batch *b, int on)
{
    if (!b) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    const bool want = on != 0;
    if (want == b->pairs) return LBMX_OK;
    for (int slot : b->order)
        if (!b->slots[slot].pending)
            return lbmx_fail(LBMX_ERR_STATE, "the step mode can change only before lattices start");
    if (want && ((b->epoch & 1) || (b->use_graphs && (b->chunk & 1))))
        return lbmx_fail(LBMX_ERR_STATE, "pair mode needs an even epoch and an even chunk");
    b->pairs = want;
    drop_graphs(b);
    b->dirty = true;
    return LBMX_OK;
}

LBMX_API int lbmx_batch_set_chunk(lbmx_batch *b, int steps)
{
    if (!b || steps < 1 || steps > 4096) return lbmx_fail(LBMX_ERR_ARG, "chunk must be in [1, 4096]");
    if (b->pairs && (steps & 1)) return lbmx_fail(LBMX_ERR_ARG, "pair mode needs an even chunk");
    if (steps != b->chunk) drop_graphs(b);
    b->chunk = steps;
    b->use_graphs = steps > 1;
    return LBMX_OK;
}

LBMX_API int lbmx_batch_tune(lbmx_batch *b, int trial_steps, lbmx_tuning *chosen)
{
    if (!b || trial_steps < 1) return lbmx_fail(LBMX_ERR_ARG, "invalid argument");
    if (b->pairs) {   // the pair kernels have fixed shapes; the tuning is not used
        if (chosen) *chosen = b->tun;
        return LBMX_OK;
    }
    for (int slot : b->order)
        if (!b->slots[slot].pending)
            return lbmx_fail(LBMX_ERR_STATE, "tuning needs lattices that have not started");
    if (b->order.empty()) return lbmx_fail(LBMX_ERR_STATE, "tuning needs lattices");
    const lbmx_tuning keep = b->tun;
    const long long epoch0 = b->epoch;
    const bool graphs0 = b->use_graphs;
    b->use_graphs = false;
    // one tile per workgroup (order 5) at every workgroup / tile shape, and the best
    // persistent shape as a reference
    std::vector<lbmx_tuning> cand;
    for (int threads : {64, 128, 256})
        for (int ilp : {1, 2, 4, 8}) {
            const int tile = threads * ilp;
            if (tile >= 256 && tile <= 2048) cand.push_back(lbmx_tuning{threads, tile, 16, 5});
        }
    cand.push_back(lbmx_tuning{256, 1024, 16, 0});
    double best = 1e300;
    lbmx_tuning pick = keep;
    int rc = LBMX_OK;
    // initial states once (trial epochs start at 0); each candidate then only gets
    // its own tile table and the epoch back at 0, so its steps stay in range
    b->epoch = 0;
    b->dirty = true;
    for (const lbmx_tuning &t : cand) {
        b->tun = t;
        b->dirty = true;
        b->epoch = 0;
        rc = commit(b);
        if (!rc) rc = lbmx_batch_enqueue(b, 2);
        double ms = 0.0;
        if (!rc) rc = lbmx_batch_time(b, trial_steps, &ms);
        if (rc) break;
        if (ms < best) {
            best = ms;
            pick = t;
        }
    }
    b->use_graphs = graphs0;
    b->tun = rc ? keep : pick;
    b->epoch = epoch0;
    for (int slot : b->order) b->slots[slot].pending = 1;
    b->dirty = true;
    drop_graphs(b);
    if (rc) return rc;
    if (chosen) *chosen = pick;
    return LBMX_OK;
}

}  // extern "C"
