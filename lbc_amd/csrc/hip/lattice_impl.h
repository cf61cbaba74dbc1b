// lattice_impl.h -- internal definition of lbmx_lattice and shared host helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "lbmx.h"
#include "pull_kernels.h"
#include "step_kernels.h"
#include "tdibc_kernels.h"

namespace lbmx {
struct ChanAux;
}

constexpr int kQ = 9;

int lbmx_fail(int code, const char *fmt, ...);
int lbmx_device_info(int *ncu);
int lbmx_bc_on_column(int t);
int lbmx_bc_nvals(int t);
int lbmx_bc_check(int t, int index, int Nx, int Ny);

#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess)                                                      \
            return lbmx_fail(LBMX_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)

struct lbmx_lattice {
    lbmx_config cfg{};
    int Nx = 0, Ny = 0, pitch = 0;
    size_t plane = 0;      // elements per population plane
    size_t esize = 8;      // bytes per stored population
    hipStream_t stream = nullptr;
    void *pop[2] = {nullptr, nullptr};
    int cur = 0;           // buffer holding the state the next step starts from
    int launch_mode = 0;   // LBMX_LAUNCH_* (multi-step launches of the gather form)
    double *R = nullptr, *U = nullptr, *V = nullptr;

    // geometry
    int geom = 0;
    uint8_t *cls = nullptr;
    double *lut = nullptr;
    double *om_arr = nullptr;
    double om_uniform = 1.0;
    long long nsolid = 0;
    lbmx::Link *links = nullptr;
    int nlinks = 0;
    std::vector<lbmx::Link> hlinks;          // host copy of the link list
    std::vector<long long> link_target;      // each link's solid node, y * Nx + x
    double *partials = nullptr;
    int npartials = 0;
    double *force_dev = nullptr;
    double *force_host = nullptr;
    // force record {t, L, D} per step (lbmx_set_force_record)
    double *frec = nullptr;
    int frec_cap = 0;
    long long frec_row = 0;
    double frec_dt = 1.0, frec_cf = 1.0;
    int edge_rows = 0, edge_cols = 0;

    // boundary table (as given) and the device program compiled from it
    std::vector<int> types, index, vals_set;
    std::vector<double *> val1, val2;
    lbmx::BcDesc *bc_dev = nullptr;
    int nprog = 0;
    std::vector<int> prog_edge;   // edge length of each program entry
    std::vector<int> prog_type, prog_index;
    std::vector<lbmx::BcDesc> prog_desc;   // host copy of bc_dev
    // host copy of the solid mask, Ny x Nx row-major (empty: no solid node)
    std::vector<uint8_t> hsolid;
    // the work after streaming as k_tail launches: program entries per launch and
    // whether the launch also takes the stale-edge job (planned by lbmx_plan_tail)
    struct TailLaunch {
        std::vector<int> entry;
        int edges = 0;
    };
    std::vector<TailLaunch> tail;
    int tail_planned = 0;

    int has_geometry = 0, has_state = 0, lag_arrays = 0;
    double Fx_lag = 0.0, Fy_lag = 0.0;
    // output moments (compute_macroscopic_fields' result); R/U/V keep the lagged
    // moments the last step read, which k_materialize needs
    double *Ro = nullptr, *Uo = nullptr, *Vo = nullptr;
    int other_is_ref = 0;        // the non-current buffer holds the reference's f
    lbmx::StepFlags last{};      // flags of the last step run
    // gather form (pull_kernels.h): buffers hold post-collision populations;
    // pop[pstar] is the one the next step streams from
    int pull = 0, pstar = 0;
    // bounded multi-step launch (chan_kernels.h, chan.hip): lattices with walls,
    // bodies or regularized columns
    int chan_ok = -1;              // lattice qualifies (-1: not checked since a change)
    int chan_lo = -1, chan_hi = -1;   // program entries of the x = 0 / x = Nx-1 columns
    int chan_aux_dirty = 1;          // chan_aux must be rewritten before the next launch
    uint2 *chan_node = nullptr;      // per node: class | link bits << 8, first link
    double *chan_lq = nullptr;       // per link: q or -1 (plain bounce-back)
    double *chan_ldf = nullptr;      // per step of a launch and link: force term
    double *chan_part = nullptr;     // per step and link block: force partials
    lbmx::ChanAux *chan_aux = nullptr;   // the kernels' rare-path tables
    unsigned *chan_done = nullptr;   // force kernel: blocks finished
    int4 *chan_work = nullptr;       // work items of a launch: x0, y0, y1, bounded path
    long long chan_work_key = -1;    // the launch shape chan_work was built for
    int chan_nwork = 0, chan_nspecial = 0;
    // forces summed inside the launches: a launch's link terms (one half of
    // chan_ldf) go through stage A in the next launch (block values into the same
    // half of chan_fpart) and stage B in the one after (totals, record rows)
    int chan_ldf_half = 0;           // half of chan_ldf the next launch writes
    double *chan_fpart = nullptr;    // two halves of per-block values
    int chan_ja = 0, chan_jb = 0;    // a launch awaits stage A (and B) / stage B
    int chan_ja_half = 0, chan_jb_half = 0;
    long long chan_ja_row = 0, chan_jb_row = 0;   // their first record rows
    std::vector<unsigned> chan_plain_rows;   // per row: prefix counts of non-plain nodes
    int nclasses = 0;                // fluid classes (geom 1)
    double hlut1 = 1.0;              // omega of class 1

    // time-domain impedance boundary controller (lbmx_set_tdibc)
    struct {
        int on = 0, bc = -1;
        lbmx::TdibcParams p{};
        lbmx::TdibcState *st = nullptr;
        double *last = nullptr;
        int *active = nullptr;
        double *hist = nullptr;
        long long step = 0;
    } tdibc;

    // density probes (lbmx_set_probes)
    struct {
        int n = 0, capacity = 0;
        long long *nodes = nullptr;
        double *hist = nullptr;
        long long step = 0;
    } probes;
    int64_t steps = 0;
    int ncu = 0, nxcd = 8, bulk_grid = 0;
    int64_t bytes = 0;
    double *staging = nullptr;
    size_t staging_bytes = 0;
    // lbmx_write_vtk: packed big-endian float32 output (16 B per node) on the
    // device and a pinned host chunk it is copied through
    uint32_t *vtk_dev = nullptr;
    uint32_t *vtk_host = nullptr;
    // lbmx_totals: partials of the fixed-order reduction, then the two totals
    double *tot_dev = nullptr;
};

lbmx::Geo lbmx_geo(const lbmx_lattice *L);
lbmx::StepFlags lbmx_flags(const lbmx_lattice *L);
int lbmx_put_pops(lbmx_lattice *L, int which, const double *aos);
int lbmx_get_pops(lbmx_lattice *L, int which, double *aos);
int lbmx_put_field(lbmx_lattice *L, double *dev, const double *host);
int lbmx_get_field(lbmx_lattice *L, double *host, const double *dev);
int lbmx_launch_moments(lbmx_lattice *L);
int lbmx_materialize_other(lbmx_lattice *L);
void lbmx_reset_run(lbmx_lattice *L);
// Returns the lattice to the reference's two-buffer state if it runs in gather form.
int lbmx_ensure_push(lbmx_lattice *L);

// bounded multi-step launch (chan.hip)
void lbmx_chan_free(lbmx_lattice *L);
int lbmx_chan_eligible(lbmx_lattice *L);
int lbmx_chan_fits(const lbmx_lattice *L);
int lbmx_chan_run(lbmx_lattice *L);
int lbmx_chan_flush(lbmx_lattice *L);
int lbmx_chan_depth();
