// batch_impl.h -- internal definition of lbmx_batch (shared by batch.hip and server.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <vector>

#include "batch_kernels.h"
#include "lattice_impl.h"
#include "totals_kernels.h"
#include "lbmx_batch.h"

constexpr int kBatchMaxSlots = 8192;

// First-fit device arena with coalescing free list; one block per lattice.
struct BatchArena {
    char *base = nullptr;
    size_t cap = 0, used = 0, peak = 0;
    std::map<size_t, size_t> free_;   // offset -> bytes
    int init(size_t bytes);
    void *alloc(size_t bytes);
    void release(void *p, size_t bytes);
    void destroy();
};

// Byte offsets of one lattice's arrays inside its arena block.
struct BatchLayout {
    size_t buf[2], fbuf, om, out[3], saved[2], state, probe, init_a, init_b, total;
    unsigned plane, fplane;
    int gap;
};

struct BatchStaging {
    void *p = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
};

struct BatchSlot {
    int used = 0;
    int pending = 0;          // waiting for its first step (descriptor not on the device yet)
    int fields_set = 0;
    lbmx_case c{};
    BatchLayout lay{};
    char *block = nullptr;
    lbmx::BatchLat d{};
    lbmx::BatchInit init{};
};

struct lbmx_batch {
    int precision = 0;
    size_t esize = 8;
    hipStream_t stream = nullptr;
    int ncu = 256;
    BatchArena arena;
    lbmx::BatchLat *lats_dev = nullptr;
    lbmx::Tile *tiles_dev = nullptr;
    long long tiles_cap = 0;
    int n_interior = 0;   // interior tiles of the committed table
    int grid5 = 0;        // interior grid of tile order 5 (n_interior rounded up)
    lbmx::BatchCtl *ctl_dev = nullptr;
    std::vector<BatchSlot> slots;
    std::vector<int> order;          // lattices in tile-table order
    std::vector<lbmx::Tile> tile_host;
    long long epoch = 0;             // host mirror of ctl->epoch (steps enqueued)
    bool dirty = true;
    lbmx_tuning tun{256, 512, 16, 5};
    int chunk = 32;
    bool use_graphs = true;
    // pair mode: two steps per pass (k_pair_band / k_pair_core); the epoch then
    // advances in steps of two
    bool pairs = false;
    int grid_band = 0;   // band-tile grid of pair mode (rounded like grid5)
    std::map<int, hipGraphExec_t> graphs;
    // graphs replaced while launches of them may still be queued on the stream:
    // destroyed once the stream is known to be idle
    std::vector<hipGraphExec_t> retired;
    std::vector<BatchStaging> ring;
    int ring_next = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // lbmx_batch_spectrum's buffers (allocated on first use)
    double *spec_dev = nullptr, *spec_host = nullptr;
    // divergence guard word of every descriptor slot (BatchLat::status)
    unsigned *status_dev = nullptr;
    // lbmx_batch_totals: partials of the fixed-order reduction, then the two totals
    double *tot_dev = nullptr;
};

int lbmx_batch_enqueue(lbmx_batch *b, int nsteps);
const BatchSlot *lbmx_batch_slot(const lbmx_batch *b, int id);
// Checks a spectrum request against the case's probes and length.
int lbmx_tl_check(const lbmx_case *c, const lbmx_tl_spec *tl);
// Probe spectra of a lattice's history (k_spectrum on `stream`, through the device
// buffer dev and the pinned host buffer host, 4 * 64 doubles each) and the TL.
int lbmx_batch_spectrum_of(const BatchSlot &S, const lbmx_tl_spec &tl, hipStream_t stream,
                           double *dev, double *host, lbmx_result *out);
// FP64 mass and kinetic energy of a completed keep_moments lattice into out[0..1]
// (device), partials in `scratch` (2 * kTotMaxBlocks doubles), on `stream`.
int lbmx_batch_totals_of(const lbmx_batch *b, int id, double *scratch, double *out,
                         hipStream_t stream);
