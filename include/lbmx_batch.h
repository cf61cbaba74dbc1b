/*
 * lbmx_batch.h -- many independent lattices advanced together on one GPU, and a
 * request server built on them.
 *
 * The reference runs one lattice per process (cases/<case>/<case>.c main loop,
 * one main_lbm call per step, src/core/main_lbm.c:5-62).  This header adds what
 * that design cannot express and has no counterpart in the reference: a batch
 * of lattices of uneven sizes whose time steps run as one launch pair per step
 * over all of them (captured in a HIP graph), and a server that admits solve requests into
 * such a batch while it runs (continuous batching).
 *
 * Each lattice is described by an lbmx_case: a program made of the reference's
 * own building blocks, applied in the reference's order within a step
 * (main_lbm.c:31-61):
 *   collision      BGK + Guo body force (collision.c:24-37), lagged moments
 *   walls          solid rows/columns with half-way bounce-back (bounce_back.c)
 *   streaming      pull with periodic wrap (streaming.c:21-42)
 *   boundary table, in this order when present:
 *     convective_inlet(0), convective_outlet(Nx-1)          (convective_*.c)
 *     velocity_bottom_regularized(0) over the bottom-wall opening, its normal
 *       velocity from a per-node time-domain impedance ODE   (cases/nit/nit.c:13-47)
 *     velocity_top(Ny-1)                                     (velocity_top.c)
 *   moments        compute_macroscopic_fields.c:19-48
 * A case run in a batch gives, in LBMX_F64, bit for bit the result of the same
 * program run by the reference (and by lbmx_lattice); in LBMX_F32_SHIFTED and
 * LBMX_F32_WIDE, bit for bit the result of an lbmx_lattice of that precision.
 * Results never depend
 * on the other lattices in the batch, on the launch configuration or on the
 * HIP-graph chunking.
 *
 * Same error convention as lbmx.h (LBMX_OK or a negative code, lbmx_last_error).
 */
#ifndef LBMX_BATCH_H
#define LBMX_BATCH_H

#include <stddef.h>

#include "lbmx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* lbmx_case.walls: solid rows / columns (the mask of the reference drivers). */
#define LBMX_WALL_BOTTOM 1u   /* row 0 (except the opening gap_x0..gap_x1) */
#define LBMX_WALL_TOP 2u      /* row Ny-1 */
#define LBMX_WALL_LEFT 4u     /* column 0 */
#define LBMX_WALL_RIGHT 8u    /* column Nx-1 */

/* lbmx_case.init: the state the first step starts from (both buffers hold the
 * equilibrium of these moments; the first collision reads them as its lagged
 * moment arrays, collision.c:24-32). */
#define LBMX_INIT_PULSE 0     /* rho = rho0 + amp*exp(-ln2*((x-x0)/halfwidth)^2), u = v = 0 */
#define LBMX_INIT_TGV 1       /* Taylor-Green vortex of lbmx_init_taylor_green with
                                 kx = 2pi/Nx, ky = 2pi/Ny */
#define LBMX_INIT_FIELDS 2    /* rho/u/v arrays given to lbmx_batch_set_fields */

/* Multipole admittance of the impedance boundary (cases/nit/nit.c:13-47,
 * impedance_*Pa.txt:1-7): two real poles (A, lambdas), two complex pairs
 * (B, C, alpha, beta), and Zinf. */
typedef struct {
    double Zinf, A[2], lambdas[2], B[2], C[2], alpha[2], beta[2];
} lbmx_impedance;

typedef struct {
    int Nx, Ny;                /* 3 <= Ny, 3 <= Nx, Nx*Ny*9 < 2^31 */
    int per_x, per_y;          /* periodic axes */
    unsigned walls;            /* LBMX_WALL_* */
    int gap_x0, gap_x1;        /* opening of the bottom wall (inclusive); none if gap_x1 < gap_x0 */
    double tau;                /* relaxation time */
    int sponge;                /* >0: tau ramps linearly to tau_sponge over `sponge`
                                  columns at both ends (the liner duct's absorbing layers) */
    double tau_sponge;
    double Fx, Fy;             /* Guo body force */
    int top_velocity;          /* velocity_top(Ny-1) with val1 = U_top, val2 = V_top */
    double U_top, V_top;
    int convective_ends;       /* convective_inlet(0) and convective_outlet(Nx-1) */
    int liner;                 /* impedance wall on the opening: velocity_bottom_regularized
                                  with val1 = 0 and val2 = gain * v(p'), v from the ODE of
                                  `imp` driven by p' = cs^2 (rho_lagged - rho0) per node */
    double gain;
    lbmx_impedance imp;
    int init;                  /* LBMX_INIT_* */
    double rho0;               /* reference density (and FP32 shift) */
    double amp, x0, halfwidth; /* LBMX_INIT_PULSE */
    double u0;                 /* LBMX_INIT_TGV */
    int nprobe;                /* <= 4 density probes: lagged rho before every step */
    int probe_x[4], probe_y[4];
    int steps;                 /* length of the solve */
    int keep_moments;          /* store rho/u/v when the solve completes */
} lbmx_case;

/* Launch configuration of the interior kernel (chosen by lbmx_batch_tune; default
 * {256, 512, 16, 5}). */
typedef struct {
    int threads;               /* threads per workgroup: 64, 128 or 256 */
    int tile;                  /* nodes per tile: threads * {1 .. 16}, >= 256 */
    int groups_per_cu;         /* persistent workgroups per compute unit (orders 0-4) */
    int order;                 /* tiles to workgroups: 5 one tile per workgroup (grid
                                  covering the table); persistent grids: 0 interleaved
                                  (tile t on workgroup t mod G), 1 one contiguous band per
                                  XCD, 2 one contiguous range per workgroup, 3 / 4 bands of
                                  4 / 16 tiles dealt to one XCD each */
} lbmx_tuning;

typedef struct lbmx_batch lbmx_batch;

/* precision: LBMX_F64, LBMX_F32_SHIFTED or LBMX_F32_WIDE.  budget: bytes of device memory the
 * batch may hold for its lattices (one arena, allocated here). */
LBMX_API int lbmx_batch_create(int precision, size_t budget, lbmx_batch **out);
LBMX_API void lbmx_batch_destroy(lbmx_batch *b);

/* Device bytes one case occupies in a batch of this precision. */
LBMX_API size_t lbmx_case_bytes(int precision, const lbmx_case *c);
/* Checks a case against the rules lbmx_batch_add enforces (no device calls). */
LBMX_API int lbmx_case_check(const lbmx_case *c);

/* Adds a lattice; it starts stepping with the next lbmx_batch_run.
 * LBMX_ERR_NOMEM when the budget cannot hold it (nothing is changed then). */
LBMX_API int lbmx_batch_add(lbmx_batch *b, const lbmx_case *c, int *id);
/* Initial moments [Ny][Nx] for an LBMX_INIT_FIELDS case, before its first run. */
LBMX_API int lbmx_batch_set_fields(lbmx_batch *b, int id, const double *rho, const double *u,
                                   const double *v);
/* Releases a lattice and its memory. */
LBMX_API int lbmx_batch_remove(lbmx_batch *b, int id);

/* Advances every lattice nsteps steps (a lattice stops at its own `steps`).
 * Asynchronous; whole chunks run as one HIP-graph launch. */
LBMX_API int lbmx_batch_run(lbmx_batch *b, int nsteps);
/* Same, timed with HIP events on the batch's stream (device milliseconds). */
LBMX_API int lbmx_batch_time(lbmx_batch *b, int nsteps, double *ms);
LBMX_API int lbmx_batch_synchronize(lbmx_batch *b);

/* Steps completed by lattice id (as enqueued; synchronize before reading results). */
LBMX_API int lbmx_batch_steps_done(lbmx_batch *b, int id, int *steps);
/* rho/u/v [Ny][Nx] at completion (keep_moments); NaN on solid nodes. */
LBMX_API int lbmx_batch_moments(lbmx_batch *b, int id, double *rho, double *u, double *v);
/* Probe history [(steps+1)][nprobe]: row t = lagged rho before step t. Returns rows. */
LBMX_API int lbmx_batch_probes(lbmx_batch *b, int id, double *hist, int capacity);

/* Divergence guard.  Every collision's moments (the state after the previous
 * step) are checked at every fluid node; a violation marks the lattice (the word
 * is the union of the kinds met over its nodes and steps) while the lattices
 * beside it go on unchanged -- the reference's drivers instead end the whole run
 * (cases/poiseuille/poiseuille.c:139-142).  A marked lattice's results are not
 * meaningful; the server stops such a request at the next chunk boundary. */
#define LBMX_GUARD_NAN 1u     /* a NaN density or velocity */
#define LBMX_GUARD_RHO 2u     /* density <= 0 (the moments floor it at 1e-16; flagged
                                 at or below 1e-12) */
#define LBMX_GUARD_SPEED 4u   /* |u| or |v| above 0.2 (lattice units) */
/* *guard = the LBMX_GUARD_* bits lattice id has tripped so far (0: healthy). */
LBMX_API int lbmx_batch_status(lbmx_batch *b, int id, unsigned *guard);

/* FP64 totals of a completed keep_moments lattice's final moments over its fluid
 * nodes: mass = sum rho, energy = sum rho (u^2 + v^2) / 2, in a fixed summation
 * order (the result depends only on the field). */
LBMX_API int lbmx_batch_totals(lbmx_batch *b, int id, double *mass, double *energy);

/* Times candidate launch configurations on the batch as it stands (every lattice
 * must still be waiting for its first step: trial steps run on scratch state and
 * the lattices are initialised afterwards) and keeps the fastest. */
LBMX_API int lbmx_batch_tune(lbmx_batch *b, int trial_steps, lbmx_tuning *chosen);
LBMX_API int lbmx_batch_set_tuning(lbmx_batch *b, const lbmx_tuning *t);
LBMX_API int lbmx_batch_get_tuning(lbmx_batch *b, lbmx_tuning *t);
/* Steps per HIP-graph launch (default 32). */
LBMX_API int lbmx_batch_set_chunk(lbmx_batch *b, int steps);

/* Pair mode (on != 0): two time steps per pass over each lattice.  A lattice's
 * core (nodes at least 3 from the top and bottom rows, and from the side columns
 * unless every side node wraps, is a wall or is a convective end) takes both
 * steps in one launch, its intermediate step kept in LDS; the other nodes, where
 * the remaining boundary rules live, take one step per launch through a small
 * frame buffer.  Results
 * are bit-identical to one step per launch; HBM traffic per step about halves.
 * The batch then advances in steps of two (lbmx_batch_run, lbmx_batch_time and
 * the chunk must be even; a lattice with an odd `steps` takes a single step at
 * the end), launch shapes are fixed (lbmx_batch_tune returns the current tuning)
 * and the mode can change only before any lattice has started.  Off by default;
 * a server with an even chunk uses it. */
LBMX_API int lbmx_batch_set_pairs(lbmx_batch *b, int on);

/* ---- server: continuous batching of solve requests ------------------------------
 * A worker thread owns a batch.  Submitted requests wait in arrival order and are
 * admitted whenever the memory budget can hold them; between graph chunks the
 * finished lattices leave and the waiting ones join, so requests that arrive
 * while others run overlap with them on the device.  A finished request's probe
 * histories are reduced on the device to spectra (lbmx_tl_spec), its final
 * moments to FP64 mass and energy, and its memory is released.  A request whose
 * divergence guard trips is stopped at the next chunk boundary with status
 * LBMX_ERR_DIVERGED; the requests beside it are not affected. */
typedef struct {
    int nfreq;                 /* <= 64 frequencies, linear from f_lo to f_hi (cycles/step) */
    double f_lo, f_hi;
    int up_probe, down_probe;  /* probe indices of the case */
    int up_t0, up_t1;          /* samples [t0, t1) of the upstream probe (incident pulse) */
    int down_t0, down_t1;      /* samples [t0, t1) of the downstream probe (transmitted) */
} lbmx_tl_spec;

typedef struct {
    long long tag;             /* as submitted */
    int status;                /* LBMX_OK, LBMX_ERR_DIVERGED (see guard) or another error */
    int steps;
    double t_submit, t_admit, t_done;   /* seconds on the server's monotonic clock */
    int nfreq;
    double tl_db[64];          /* 20 log10(|P_up(f)| / |P_down(f)|) */
    double up_abs[64], down_abs[64];
    unsigned guard;            /* LBMX_GUARD_* bits (status LBMX_ERR_DIVERGED) */
    double mass, energy;       /* FP64 totals of the final state (lbmx_batch_totals) */
} lbmx_result;

typedef struct lbmx_server lbmx_server;

/* precision/budget as for a batch; chunk = steps per graph launch (an even chunk
 * runs the batch in pair mode, lbmx_batch_set_pairs). */
LBMX_API int lbmx_server_create(int precision, size_t budget, int chunk,
                                const lbmx_tuning *tuning, lbmx_server **out);
/* Queues one request (thread-safe, returns at once).  tl may be NULL: a request
 * without a spectrum (a channel solve) returns its status, timing and the FP64
 * mass and energy of its final state, with nfreq = 0. */
LBMX_API int lbmx_server_submit(lbmx_server *s, const lbmx_case *c, const lbmx_tl_spec *tl,
                                long long tag);
/* Copies up to max finished results (oldest first); returns the count. */
LBMX_API int lbmx_server_poll(lbmx_server *s, lbmx_result *out, int max);
/* Seconds on the server's clock (the clock of t_submit / t_done). */
LBMX_API double lbmx_server_clock(lbmx_server *s);
/* Counters: requests waiting, running, finished; peak bytes held. */
LBMX_API int lbmx_server_stats(lbmx_server *s, int *waiting, int *running, int *finished,
                               size_t *peak_bytes);
/* Stops the worker after the running and waiting requests finish, then frees. */
LBMX_API void lbmx_server_destroy(lbmx_server *s);

/* The server's spectrum and TL (the same device reduction, k_spectrum) of a
 * completed lattice of a batch: out receives nfreq, tl_db, up_abs and down_abs
 * (tag -1, times 0).  A request's result from the server equals this for the
 * same case run alone in a batch. */
LBMX_API int lbmx_batch_spectrum(lbmx_batch *b, int id, const lbmx_tl_spec *tl,
                                 lbmx_result *out);

#ifdef __cplusplus
}
#endif

#endif
