// chan.hip -- host side of the bounded multi-step launch (chan_kernels.h): which
// lattices qualify, the per-node tables the kernels read, and one launch of
// several steps.  lbmx_step uses it for multi-step calls on lattices with walls,
// bodies or regularized inlet/outlet columns; every call still ends with at least
// one step of the one-step path, so between calls both buffers hold exactly what
// that path leaves (the launch itself leaves the non-current buffer at the start
// state of the launch, not at the reference's f_new).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "chan_kernels.h"
#include "lattice_impl.h"

using namespace lbmx;

namespace {

constexpr int kChanDepth = 3;    // steps per k_chan launch (levels: kChanDepth + 1)
constexpr int kChanLW = 64;      // lanes per level
static_assert(kChanDepth <= kChanMaxLev, "k_chan_force sums at most kChanMaxLev steps");

int env_int(const char *name, int dflt)
{
    const char *v = getenv(name);
    return v ? atoi(v) : dflt;
}

}   // namespace

// doubles of one half of chan_fpart: per step, pair and link block two values
static size_t chan_fpart_half()
{
    return (size_t)kChanDepth * 2 * kChanFusedBlocks * 2;
}

static double *chan_ldf_at(const lbmx_lattice *L, int half)
{
    return L->chan_ldf ? L->chan_ldf + (size_t)half * kChanDepth * L->nlinks : nullptr;
}

void lbmx_chan_free(lbmx_lattice *L)
{
    hipFree(L->chan_node);
    hipFree(L->chan_lq);
    hipFree(L->chan_ldf);
    hipFree(L->chan_part);
    hipFree(L->chan_aux);
    hipFree(L->chan_done);
    hipFree(L->chan_work);
    L->chan_work = nullptr;
    L->chan_work_key = -1;
    L->chan_plain_rows.clear();

    L->chan_node = nullptr;
    L->chan_lq = nullptr;
    L->chan_ldf = nullptr;
    L->chan_part = nullptr;
    L->chan_aux = nullptr;
    L->chan_done = nullptr;
    L->chan_ok = -1;
    hipFree(L->chan_fpart);
    L->chan_fpart = nullptr;
    L->chan_ja = L->chan_jb = 0;
    L->chan_ldf_half = 0;
}

// The geometry and boundary table allow the bounded launch: every program entry
// is a regularized column at x = 0 (velocity or pressure inlet) or x = Nx-1
// (outlet) on a closed x axis, no fluid node keeps a stale entry (streaming
// sources outside the domain only on those columns, which rewrite all nine
// populations), each column node's interior neighbour is fluid, and no
// interpolated link with q <= 1/2 has a solid behind node.
static int chan_static_check(lbmx_lattice *L)
{
    const int Nx = L->Nx, Ny = L->Ny, px = L->cfg.per_x, py = L->cfg.per_y;
    L->chan_lo = L->chan_hi = -1;
    for (int b = 0; b < L->nprog; b++) {
        const int t = L->prog_type[b], idx = L->prog_index[b];
        if ((t == LBMX_BC_VELOCITY_INLET_REG || t == LBMX_BC_PRESSURE_INLET_REG) && idx == 0 &&
            L->chan_lo < 0)
            L->chan_lo = b;
        else if ((t == LBMX_BC_VELOCITY_OUTLET_REG || t == LBMX_BC_PRESSURE_OUTLET_REG) &&
                 idx == Nx - 1 && L->chan_hi < 0)
            L->chan_hi = b;
        else
            return 0;
    }
    if ((L->chan_lo >= 0 || L->chan_hi >= 0) && px) return 0;
    static const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    auto solid = [&](int x, int y) {
        return !L->hsolid.empty() && L->hsolid[(size_t)y * Nx + x];
    };
    // only nodes on a closed edge can have a source outside the domain
    auto edge_ok = [&](int x, int y) -> bool {
        if (solid(x, y)) return true;
        const bool col = (L->chan_lo >= 0 && x == 0) || (L->chan_hi >= 0 && x == Nx - 1);
        if (col) return !solid(x == 0 ? 1 : Nx - 2, y);
        for (int k = 1; k < 9; k++) {
            const int sx = x - CX[k], sy = y - CY[k];
            if ((!px && (sx < 0 || sx >= Nx)) || (!py && (sy < 0 || sy >= Ny))) return false;
        }
        return true;
    };
    if (!py)
        for (int x = 0; x < Nx; x++)
            if (!edge_ok(x, 0) || !edge_ok(x, Ny - 1)) return 0;
    if (!px)
        for (int y = 0; y < Ny; y++)
            if (!edge_ok(0, y) || !edge_ok(Nx - 1, y)) return 0;
    if (L->cfg.use_ibb)
        for (const Link &l : L->hlinks)
            if (l.nb >= 0 && l.q <= 0.5 && l.bsolid) return 0;
    return 1;
}

int lbmx_chan_eligible(lbmx_lattice *L)
{
    static int disabled = -1;
    if (disabled < 0) disabled = getenv("LBMX_NO_CHAN") ? 1 : 0;
    if (disabled || !L->has_geometry || L->tdibc.on || L->probes.n || L->pull ||
        (L->cfg.flags & LBMX_EXACT_STATE))
        return 0;
    // periodic all-fluid lattices without a table run the plain gather form
    if (L->geom == 0 && L->cfg.per_x && L->cfg.per_y && L->nprog == 0) return 0;
    if (L->chan_ok < 0) {
        L->chan_ok = chan_static_check(L);
        L->chan_aux_dirty = 1;
    }
    return L->chan_ok;
}

// The launch pays once the lattice gives it enough workgroups (as k_stream).
int lbmx_chan_fits(const lbmx_lattice *L)
{
    const int tx = kChanLW - 2 * kChanDepth;
    const long long strips = (L->Nx + 1 + tx - 1) / tx;
    return strips * ((L->Ny + 31) / 32) >= 200;
}

int lbmx_chan_depth() { return kChanDepth; }

// Per-node class, link bits and first link; link q's (once per geometry).
static int chan_prepare(lbmx_lattice *L)
{
    if (L->chan_node) return LBMX_OK;
    const int Nx = L->Nx, Ny = L->Ny, pitch = L->pitch;
    const size_t nn = (size_t)Ny * pitch;
    std::vector<uint2> node(nn, make_uint2(0u, 0u));
    if (L->geom == 0) {
        for (int y = 0; y < Ny; y++)
            for (int x = 0; x < Nx; x++) node[(size_t)y * pitch + x].x = 1;
    } else {
        std::vector<uint8_t> cls(nn);
        HIPCHK(hipStreamSynchronize(L->stream));
        HIPCHK(hipMemcpy(cls.data(), L->cls, nn, hipMemcpyDeviceToHost));
        for (size_t n = 0; n < nn; n++) node[n].x = cls[n];
    }
    std::vector<double> lq(L->hlinks.size());
    for (size_t l = 0; l < L->hlinks.size(); l++) {
        const Link &k = L->hlinks[l];
        const size_t n = (size_t)k.n;
        if ((node[n].x >> 8) == 0) node[n].y = (unsigned)l;
        node[n].x |= 1u << (7 + k.k);   // directions 1..8 -> bits 8..15
        const bool plain = !L->cfg.use_ibb || k.nb < 0 || (k.q > 0.5 && k.bsolid);
        lq[l] = plain ? -1.0 : k.q;
    }
    const int nblk = L->npartials;
    if (hipMalloc((void **)&L->chan_node, nn * sizeof(uint2)) != hipSuccess ||
        hipMalloc((void **)&L->chan_aux, sizeof(ChanAux)) != hipSuccess ||
        hipMalloc((void **)&L->chan_done, sizeof(unsigned)) != hipSuccess ||
        hipMalloc((void **)&L->chan_fpart, chan_fpart_half() * 2 * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&L->chan_part, (size_t)kChanDepth * (nblk > 0 ? nblk : 1) *
                                              kPartial * sizeof(double)) != hipSuccess) {
        lbmx_chan_free(L);
        return lbmx_fail(LBMX_ERR_NOMEM, "bounded launch tables");
    }
    if (!lq.empty() &&
        (hipMalloc((void **)&L->chan_lq, lq.size() * sizeof(double)) != hipSuccess ||
         hipMalloc((void **)&L->chan_ldf, 2 * (size_t)kChanDepth * lq.size() * sizeof(double)) !=
             hipSuccess)) {
        lbmx_chan_free(L);
        return lbmx_fail(LBMX_ERR_NOMEM, "bounded launch link tables");
    }
    HIPCHK(hipMemcpy(L->chan_node, node.data(), nn * sizeof(uint2), hipMemcpyHostToDevice));
    // per row, the prefix counts of the nodes a plain workgroup cannot compute:
    // solid, with links, on a column of the table, or on a closed edge of the domain
    {
        const bool uni = L->geom == 0 || (L->geom == 1 && L->nclasses == 1);
        L->chan_plain_rows.assign((size_t)Ny * (Nx + 1), 0u);
        for (int y = 0; y < Ny; y++) {
            unsigned *pre = &L->chan_plain_rows[(size_t)y * (Nx + 1)];
            for (int x = 0; x < Nx; x++) {
                const uint2 w = node[(size_t)y * pitch + x];
                const bool bad = !uni || (w.x & 0xffu) == 0u || (w.x >> 8) != 0u ||
                                 (L->chan_lo >= 0 && x == 0) ||
                                 (L->chan_hi >= 0 && x == Nx - 1) ||
                                 (!L->cfg.per_x && (x == 0 || x == Nx - 1)) ||
                                 (!L->cfg.per_y && (y == 0 || y == Ny - 1));
                pre[x + 1] = pre[x] + (bad ? 1u : 0u);
            }
        }
    }
    if (!lq.empty())
        HIPCHK(hipMemcpy(L->chan_lq, lq.data(), lq.size() * sizeof(double),
                         hipMemcpyHostToDevice));
    HIPCHK(hipMemset(L->chan_done, 0, sizeof(unsigned)));
    L->chan_aux_dirty = 1;
    return LBMX_OK;
}

// The rare-path tables (boundary value arrays included, which lbmx_set_boundaries
// reallocates).
static int chan_write_aux(lbmx_lattice *L)
{
    ChanAux x{};
    x.lq = L->chan_lq;
    x.nlinks = L->nlinks;
    if (L->chan_lo >= 0) {
        const BcDesc &d = L->prog_desc[L->chan_lo];
        x.lo_type = d.type;
        x.lo_v1 = d.v1;
        x.lo_v2 = d.v2;
    }
    if (L->chan_hi >= 0) {
        const BcDesc &d = L->prog_desc[L->chan_hi];
        x.hi_type = d.type;
        x.hi_v1 = d.v1;
        x.hi_v2 = d.v2;
    }
    x.geom = L->geom;
    x.lut = L->lut;
    x.om_arr = L->om_arr;
    x.R = L->R;
    x.U = L->U;
    x.V = L->V;
    HIPCHK(hipStreamSynchronize(L->stream));
    HIPCHK(hipMemcpy(L->chan_aux, &x, sizeof(x), hipMemcpyHostToDevice));
    L->chan_aux_dirty = 0;
    return LBMX_OK;
}

static ChanArgs chan_args(const lbmx_lattice *L, const StepFlags &s)
{
    ChanArgs a{};
    a.Nx = L->Nx;
    a.Ny = L->Ny;
    a.pitch = L->pitch;
    a.per_x = L->cfg.per_x;
    a.per_y = L->cfg.per_y;
    a.plane = L->plane;
    a.rho0 = L->cfg.rho0;
    a.wide = L->cfg.precision == LBMX_F32_WIDE;
    a.uni = L->geom == 0 || (L->geom == 1 && L->nclasses == 1);
    a.om = L->geom == 0 ? L->om_uniform : L->hlut1;
    a.force = s.force;
    a.Fx = s.Fx;
    a.Fy = s.Fy;
    a.Fx_lag = s.Fx_lag;
    a.Fy_lag = s.Fy_lag;
    a.lag = s.lag_arrays;
    a.lo = L->chan_lo >= 0;
    a.hi = L->chan_hi >= 0;
    a.node = L->chan_node;
    a.aux = L->chan_aux;
    return a;
}

// The forces of a launch's steps (terms in ldf, record rows from row0 on).  With
// at most kChanFusedBlocks link blocks the next two launches of the call sum them
// in their force workgroups (stages A and B), and k_chan_force sums those of the
// call's last two launches when it leaves the bounded launches (lbmx_chan_flush);
// with more, k_chan_force runs after every launch.
static ChanForceArgs chan_force_args(const lbmx_lattice *L, const double *ldf, int nlev,
                                     long long row0)
{
    ChanForceArgs f{};
    f.links = L->links;
    f.nlinks = L->nlinks;
    f.ldf = ldf;
    f.nlev = nlev;
    f.partials = L->chan_part;
    f.nblk = L->nlinks ? L->npartials : 0;
    f.done = L->chan_done;
    f.force_out = L->force_dev;
    f.record = L->frec;
    f.row0 = row0;
    f.capacity = L->frec_cap;
    f.dt = L->frec_dt;
    f.cf = L->frec_cf;
    return f;
}

static int chan_force_kernel(lbmx_lattice *L, const ChanForceArgs &f)
{
    hipLaunchKernelGGL(k_chan_force, dim3(f.nblk > 0 ? f.nblk : 1), dim3(kLinkThreads), 0,
                       L->stream, f);
    HIPCHK(hipGetLastError());
    return LBMX_OK;
}

int lbmx_chan_flush(lbmx_lattice *L)
{
    // the older launch first: force_out ends with the newest step
    if (L->chan_jb) {
        L->chan_jb = 0;
        const ChanForceArgs f =
            chan_force_args(L, chan_ldf_at(L, L->chan_jb_half), kChanDepth, L->chan_jb_row);
        if (int rc = chan_force_kernel(L, f)) return rc;
    }
    if (L->chan_ja) {
        L->chan_ja = 0;
        const ChanForceArgs f =
            chan_force_args(L, chan_ldf_at(L, L->chan_ja_half), kChanDepth, L->chan_ja_row);
        if (int rc = chan_force_kernel(L, f)) return rc;
    }
    return LBMX_OK;
}

// The work list of a launch: strips of TX output columns (the first starting at
// -xshift), cut into row segments; a segment whose level-1 region (columns
// x0 - (NL-1) .. x0 + TX + NL - 2, its rows +- (NL-1), wrapped on periodic axes)
// holds a node a plain workgroup cannot compute takes the bounded path and is cut
// `split` times shorter, since that path costs more per row.  Rows per segment:
// the fewest (not under 8) that keep the list within one round of the workgroups
// the CUs hold at once; a second, partial round doubles the launch's critical
// path.  Lattices too large for one round take 32-row segments.
static int chan_work(lbmx_lattice *L, int NL, int TX, int per_cu, ChanArgs &a)
{
    const int Nx = L->Nx, Ny = L->Ny;
    const int xshift = (a.hi && (Nx - 1) % TX == 0) ? 1 : 0;   // outlet shares a strip
    const int strips = (Nx + xshift + TX - 1) / TX;
    static const int rows_env = env_int("LBMX_CHAN_ROWS", 0);
    // bounded segments 4 times shorter (cylinder 2048 x 512 FP64, us per step, split
    // 1 / 2 / 4: 31.4 / 25.9 / 23.1)
    static const int split_env = env_int("LBMX_CHAN_SPLIT", 4);
    const int split = std::max(1, split_env);
    const long long key = ((long long)rows_env << 40) ^ ((long long)TX << 20) ^
                          ((long long)split << 10) ^ NL;
    if (L->chan_work && L->chan_work_key == key) {
        a.nwork = L->chan_nwork;
        return LBMX_OK;
    }
    auto row_bad = [&](int y, int xa, int xb) -> bool {   // columns [xa, xb)
        if (!L->cfg.per_y && (y < 0 || y >= Ny)) return true;
        y = ((y % Ny) + Ny) % Ny;
        const unsigned *pre = &L->chan_plain_rows[(size_t)y * (Nx + 1)];
        if (!L->cfg.per_x && (xa < 0 || xb > Nx)) return true;
        if (xb - xa >= Nx) return pre[Nx] != 0u;
        int a0 = ((xa % Nx) + Nx) % Nx, b0 = a0 + (xb - xa);
        if (b0 <= Nx) return pre[b0] != pre[a0];
        return pre[Nx] != pre[a0] || pre[b0 - Nx] != 0u;
    };
    auto bad = [&](int x0, int y0, int y1) -> bool {
        const int xa = x0 - (NL - 1), xb = x0 + TX + (NL - 1);
        for (int y = y0 - (NL - 1); y < y1 + (NL - 1); y++)
            if (row_bad(y, xa, xb)) return true;
        return false;
    };
    auto build = [&](int rows, std::vector<int4> &list) {
        list.clear();
        for (int y0 = 0; y0 < Ny; y0 += rows) {
            const int y1 = std::min(y0 + rows, Ny);
            for (int s = 0; s < strips; s++) {
                const int x0 = s * TX - xshift;
                if (!bad(x0, y0, y1)) {
                    list.push_back(make_int4(x0, y0, y1, 0));
                    continue;
                }
                const int sub = (y1 - y0 + split - 1) / split;
                for (int ya = y0; ya < y1; ya += sub) {
                    const int yb = std::min(ya + sub, y1);
                    list.push_back(make_int4(x0, ya, yb, bad(x0, ya, yb) ? 1 : 0));
                }
            }
        }
    };
    // the launch's force workgroups take slots too
    const long long slots =
        (long long)per_cu * (L->ncu > 0 ? L->ncu : 256) - 2 * kChanDepth * kChanSplit;
    std::vector<int4> list;
    if (rows_env > 0) {
        build(std::min(rows_env, Ny), list);
    } else if ((long long)strips * ((Ny + 31) / 32) > 2 * slots) {
        build(std::min(32, Ny), list);
    } else {
        int rows = 8;
        for (;; rows++) {
            build(std::min(rows, Ny), list);
            if ((long long)list.size() <= slots || rows >= Ny) break;
        }
    }
    int ns = 0;
    for (const int4 &w : list) ns += w.w;
    hipFree(L->chan_work);
    L->chan_work = nullptr;
    HIPCHK(hipMalloc((void **)&L->chan_work, list.size() * sizeof(int4)));
    HIPCHK(hipMemcpy(L->chan_work, list.data(), list.size() * sizeof(int4),
                     hipMemcpyHostToDevice));
    L->chan_work_key = key;
    L->chan_nwork = (int)list.size();
    L->chan_nspecial = ns;
    a.nwork = L->chan_nwork;
    static const int verbose = env_int("LBMX_CHAN_VERBOSE", 0);
    if (verbose)
        fprintf(stderr, "k_chan: %d workgroups per CU, %d strips, %d work items, %d bounded\n",
                per_cu, strips, L->chan_nwork, ns);
    return LBMX_OK;
}

template <typename T, bool W, int P, int R, bool NF>
static int launch_chan(lbmx_lattice *L, T *S, T *Dst, ChanArgs &a, const ChanFusedArgs &fj)
{
    constexpr int NL = kChanDepth + 1;
    constexpr int TX = ChanShape<NL, kChanLW, R>::TX;
    static int per_cu = -1;
    if (per_cu < 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, reinterpret_cast<const void *>(&k_chan<T, W, NL, kChanLW, P, R, NF>),
                kChanLW * NL, 0) != hipSuccess ||
            n < 1)
            n = 1;
        per_cu = n;
    }
    if (int rc = chan_work(L, NL, TX, per_cu, a)) return rc;
    a.work = L->chan_work;
    const unsigned nf = (fj.a_on || fj.b_on) ? 2u * kChanDepth * kChanSplit : 0u;
    const unsigned blocks = ((unsigned)a.nwork + 7) / 8 * 8 + nf;
    hipLaunchKernelGGL((k_chan<T, W, NL, kChanLW, P, R, NF>), dim3(blocks), dim3(kChanLW * NL), 0,
                       L->stream, S, Dst, a, fj);
    return LBMX_OK;
}

template <typename T>
static int chan_run_t(lbmx_lattice *L)
{
    StepFlags s = lbmx_flags(L);
    ChanArgs a = chan_args(L, s);
    T *S = (T *)L->pop[L->cur];
    T *Dst = (T *)L->pop[L->cur ^ 1];
    // this launch's link terms go to one half of ldf while its stage A reads the
    // previous launch's from the other
    a.ldf = chan_ldf_at(L, L->chan_ldf_half);
    const bool fused = L->npartials <= kChanFusedBlocks;
    ChanFusedArgs fj{};
    if (L->chan_jb) {
        fj.b_on = 1;
        fj.b_nblk = L->nlinks ? L->npartials : 0;
        fj.part_in = L->chan_fpart + L->chan_jb_half * chan_fpart_half();
        fj.force_out = L->force_dev;
        fj.record = L->frec;
        fj.row0 = L->chan_jb_row;
        fj.capacity = L->frec_cap;
        fj.dt = L->frec_dt;
        fj.cf = L->frec_cf;
    }
    if (L->chan_ja && L->nlinks) {
        fj.a_on = 1;
        fj.links = L->links;
        fj.nlinks = L->nlinks;
        fj.nblk = L->npartials;
        fj.ldf = chan_ldf_at(L, L->chan_ja_half);
        fj.part_out = L->chan_fpart + L->chan_ja_half * chan_fpart_half();
    }
    // lattices without a body force (the cylinder channel) run the force-free
    // instantiation
    const bool nf = !a.force && a.Fx == 0.0 && a.Fy == 0.0 && a.Fx_lag == 0.0 && a.Fy_lag == 0.0;
    int rc;
    if constexpr (sizeof(T) == 8)
        rc = nf ? launch_chan<T, false, 3, 3, true>(L, S, Dst, a, fj)
                : launch_chan<T, false, 3, 3, false>(L, S, Dst, a, fj);
    else if (L->cfg.precision == LBMX_F32_WIDE)
        rc = nf ? launch_chan<T, true, 4, 4, true>(L, S, Dst, a, fj)
                : launch_chan<T, true, 4, 4, false>(L, S, Dst, a, fj);
    else
        rc = nf ? launch_chan<T, false, 4, 4, true>(L, S, Dst, a, fj)
                : launch_chan<T, false, 4, 4, false>(L, S, Dst, a, fj);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    L->chan_jb = L->chan_ja;
    L->chan_jb_half = L->chan_ja_half;
    L->chan_jb_row = L->chan_ja_row;
    L->chan_ja = 0;
    if (L->nlinks || L->frec) {
        const long long row = L->frec_row;
        if (L->frec) L->frec_row += kChanDepth;
        if (fused) {
            L->chan_ja = 1;
            L->chan_ja_half = L->chan_ldf_half;
            L->chan_ja_row = row;
        } else if (int rf = chan_force_kernel(L, chan_force_args(L, a.ldf, kChanDepth, row))) {
            return rf;
        }
    }
    L->chan_ldf_half ^= 1;
    L->cur ^= 1;
    L->other_is_ref = 0;
    s.Fx_lag = L->cfg.Fx;
    s.Fy_lag = L->cfg.Fy;
    L->last = s;
    L->lag_arrays = 0;
    L->Fx_lag = L->cfg.Fx;
    L->Fy_lag = L->cfg.Fy;
    L->steps += kChanDepth;
    return LBMX_OK;
}

// One launch of lbmx_chan_depth() steps from the one-step path's state.  The
// caller calls lbmx_chan_flush after its last launch, then runs at least one step
// of the one-step path.
int lbmx_chan_run(lbmx_lattice *L)
{
    if (int rc = chan_prepare(L)) return rc;
    if (L->chan_aux_dirty)
        if (int rc = chan_write_aux(L)) return rc;
    return L->esize == 8 ? chan_run_t<double>(L) : chan_run_t<float>(L);
}
