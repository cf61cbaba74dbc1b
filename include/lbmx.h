/*
 * lbmx.h -- native, device-resident C ABI of the MI355X lattice-Boltzmann library.
 *
 * The reference library (include/lbm_core.h, lbm_bc.h) is a set of host-array
 * functions called once per time step.  This header is the thin C layer under
 * that drop-in API: a lattice lives in GPU memory between calls, its boundary
 * table is compiled once, and a time step is a short sequence of gfx950 kernels
 * on the lattice's own HIP stream.  Plain pointers and sizes only; no HIP or
 * torch types appear here, so C, ctypes or any other FFI can bind it.
 *
 * Every function returns LBMX_OK (0) or a negative error code; lbmx_last_error()
 * describes the last failure of the calling thread.
 *
 * Array conventions (the reference's, lbm_core.h:10-31):
 *   populations  double f[Ny][Nx][9]  (array of structures, k fastest)
 *   fields       double a[Ny][Nx]     (row-major)
 *   solid mask   int    m[Ny][Nx]     (nonzero = solid)
 * D2Q9 ordering cx={0,1,0,-1,0,1,-1,-1,1}, cy={0,0,1,0,-1,1,1,-1,-1},
 * w={4/9,1/9 x4,1/36 x4}, opp={0,3,4,1,2,7,8,5,6} (cases/nit/nit.c:195-202).
 */
#ifndef LBMX_H
#define LBMX_H

#include <stdint.h>

#if defined(__GNUC__)
#define LBMX_API __attribute__((visibility("default")))
#else
#define LBMX_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define LBMX_OK 0
#define LBMX_ERR_ARG (-1)     /* invalid argument or unsupported configuration */
#define LBMX_ERR_HIP (-2)     /* a HIP runtime call failed */
#define LBMX_ERR_STATE (-3)   /* call not valid in the lattice's current state */
#define LBMX_ERR_NOMEM (-4)   /* device allocation failed */
#define LBMX_ERR_DIVERGED (-5) /* a lattice's divergence guard tripped (lbmx_batch.h) */

/* Storage / arithmetic of the populations. */
#define LBMX_F64 0            /* FP64 populations, FP64 collision (reference arithmetic) */
#define LBMX_F32_SHIFTED 1    /* FP32 populations stored as f_k - w_k*rho0, FP32 collision */
#define LBMX_F32_WIDE 2       /* the same FP32 storage, collision evaluated in FP64 */

/* Boundary-condition codes: one per BCFunction of include/lbm_bc.h. */
#define LBMX_BC_CONVECTIVE_INLET 1          /* src/bc/convective_inlet.c */
#define LBMX_BC_CONVECTIVE_OUTLET 2         /* src/bc/convective_outlet.c */
#define LBMX_BC_PRESSURE_INLET 3            /* src/bc/pressure_inlet.c */
#define LBMX_BC_PRESSURE_OUTLET 4           /* src/bc/pressure_outlet.c */
#define LBMX_BC_PERIODIC_X 5                /* src/bc/periodic_x.c */
#define LBMX_BC_PERIODIC_Y 6                /* src/bc/periodic_y.c */
#define LBMX_BC_VELOCITY_INLET 7            /* src/bc/velocity_inlet.c */
#define LBMX_BC_VELOCITY_OUTLET 8           /* src/bc/velocity_outlet.c */
#define LBMX_BC_VELOCITY_TOP 9              /* src/bc/velocity_top.c */
#define LBMX_BC_VELOCITY_BOTTOM 10          /* src/bc/velocity_bottom.c */
#define LBMX_BC_VELOCITY_OUTLET_REG 11      /* src/bc/velocity_outlet_regularized.c */
#define LBMX_BC_VELOCITY_INLET_REG 12       /* src/bc/velocity_inlet_regularized.c */
#define LBMX_BC_VELOCITY_TOP_REG 13         /* src/bc/velocity_top_regularized.c */
#define LBMX_BC_VELOCITY_BOTTOM_REG 14      /* src/bc/velocity_bottom_regularized.c */
#define LBMX_BC_PRESSURE_OUTLET_REG 15      /* src/bc/pressure_outlet_regularized.c */
#define LBMX_BC_PRESSURE_INLET_REG 16       /* src/bc/pressure_inlet_regularized.c */
#define LBMX_BC_PRESSURE_TOP_REG 17         /* src/bc/pressure_top_regularized.c */
#define LBMX_BC_PRESSURE_BOTTOM_REG 18      /* src/bc/pressure_bottom_regularized.c */
#define LBMX_BC_CONVECTIVE_TOP 19           /* src/bc/convective_top.c (not declared in lbm_bc.h) */
#define LBMX_BC_CONVECTIVE_BOTTOM 20        /* src/bc/convective_bottom.c (not declared in lbm_bc.h) */
#define LBMX_BC_MAX 20

/* Lattice options (lbmx_config.flags). */
#define LBMX_EXACT_STATE 1u   /* keep the reference's full two-buffer state every step (both
                                 buffers downloadable); costs one extra population write */

typedef struct lbmx_lattice lbmx_lattice;

typedef struct {
    int Nx, Ny;
    int precision;        /* LBMX_F64, LBMX_F32_SHIFTED or LBMX_F32_WIDE */
    int per_x, per_y;     /* periodic wrap in streaming (lbm_core.h:23-24) */
    int use_ibb;          /* interpolated bounce-back at solid links (lbm_core.h:30) */
    double Fx, Fy;        /* Guo body force (lbm_core.h:25-26) */
    double rho0;          /* shift density of the FP32 storage (usually 1) */
    unsigned flags;       /* LBMX_EXACT_STATE */
} lbmx_config;

LBMX_API const char *lbmx_last_error(void);
/* Name of the device the library runs on (for logs), or "" when none. */
LBMX_API const char *lbmx_device_name(void);

/* Replaces main_lbm's per-call arrays (lbm_core.h:6-31) by a resident lattice. */
LBMX_API int lbmx_create(const lbmx_config *cfg, lbmx_lattice **out);
LBMX_API void lbmx_destroy(lbmx_lattice *lat);

/* Geometry: solid mask [Ny][Nx] (required), omega_eff [Ny][Nx] (required),
 * phi [Ny][Nx] signed distance (required when use_ibb).  Copied; the link list
 * and node classes are compiled here.  Replaces the solid_mask/omega_eff/phi
 * arguments of main_lbm (lbm_core.h:19-20, 29). */
LBMX_API int lbmx_set_geometry(lbmx_lattice *lat, const int *solid_mask, const double *omega,
                      const double *phi);

/* All-fluid geometry with one relaxation rate (no host arrays). */
LBMX_API int lbmx_set_uniform(lbmx_lattice *lat, double omega);

/* Boundary table in application order (main_lbm.c:49-58): types[b] is an
 * LBMX_BC_* code, index[b] the column (inlet/outlet) or row (top/bottom). */
LBMX_API int lbmx_set_boundaries(lbmx_lattice *lat, const int *types, const int *index, int nbc);

/* Body force of the following steps (main_lbm's Fx, Fy may change per call). */
LBMX_API int lbmx_set_force(lbmx_lattice *lat, double Fx, double Fy);

/* Values of boundary b for the next step: val1/val2 are host arrays of the edge
 * length (Ny for inlet/outlet, Nx for top/bottom) or NULL when the boundary does
 * not read them (Boundary.val1/val2, lbm_types.h:22-27). */
LBMX_API int lbmx_set_bc_values(lbmx_lattice *lat, int b, const double *val1, const double *val2);

/* Starts a run from host state in the reference layout: f is what the caller's
 * `f` holds before main_lbm, f_new what its `f_new` holds (its entries that the
 * step does not overwrite survive, as in streaming.c:36-40), and rho/u/v the
 * moment arrays the first collision reads (collision.c:24-32).  f_new may be
 * NULL (then it is a copy of f). */
LBMX_API int lbmx_upload(lbmx_lattice *lat, const double *f, const double *f_new, const double *rho,
                const double *u, const double *v);

/* Starts a run from a Taylor-Green vortex generated on the device (no host
 * arrays): u = u0 (sin kx x cos ky y, -(kx/ky) cos kx x sin ky y),
 * rho = 1 - (3 u0^2/4)(cos 2kx x + (kx/ky)^2 cos 2ky y), both buffers at
 * equilibrium, moment arrays = (rho, u, v).  For benchmarks and large runs
 * (the configuration of cases/taylorgreen/taylorgreen.c with periodic axes). */
LBMX_API int lbmx_init_taylor_green(lbmx_lattice *lat, double u0, double kx, double ky);

/* Advances nsteps time steps; each is main_lbm followed by the driver's f/f_new
 * swap.  Asynchronous on the lattice's stream. */
LBMX_API int lbmx_step(lbmx_lattice *lat, int nsteps);

/* Launch grouping of multi-step runs on lattices periodic in x and y with no
 * solid node and no other boundary (no counterpart in the reference; every mode
 * gives bit-identical results).  Such runs keep post-collision populations and
 * advance several steps per launch:
 *   LBMX_LAUNCH_AUTO    row streaming (LBMX_LAUNCH_STREAM) when the lattice gives
 *                       it enough workgroups (from about 600^2; 896^2 for
 *                       LBMX_F32_WIDE), else tiles (the default)
 *   LBMX_LAUNCH_SINGLE  one step per launch (also the default when the
 *                       environment sets LBMX_NO_TEMPORAL)
 *   LBMX_LAUNCH_TILES   three steps per launch (two with LBMX_TEMPORAL_DEPTH=2) on
 *                       64 x TY tiles with their rings in LDS
 *   LBMX_LAUNCH_STREAM  5 (FP64), 4 (FP32 storage with FP64 collision) or 8
 *                       (FP32) steps per launch, each workgroup streaming down a
 *                       column strip with one wavefront per time level (under
 *                       LBMX_LAUNCH_AUTO, FP64 takes 4 on mid-size lattices and FP32
 *                       4 below 6 M cells)
 * A run of n steps that enters this form spends its first step on an in-place
 * collision; the remaining steps go in launches of the mode's depth, the last
 * few in shallower ones.
 * Lattices with walls, bodies ((interpolated) bounce-back) or regularized
 * velocity/pressure columns at x = 0 and x = Nx - 1, and no entry that stays stale,
 * run three steps per launch under LBMX_LAUNCH_STREAM, and under LBMX_LAUNCH_AUTO
 * when large enough (about 200 workgroups of 58 columns x 32 rows); every call of
 * theirs ends with 1..3 steps of one step per launch, so between calls both
 * buffers hold the reference's fields.  LBMX_LAUNCH_SINGLE and _TILES keep them
 * at one step per launch.  Other lattices ignore the mode. */
#define LBMX_LAUNCH_AUTO 0
#define LBMX_LAUNCH_SINGLE 1
#define LBMX_LAUNCH_TILES 2
#define LBMX_LAUNCH_STREAM 3
LBMX_API int lbmx_set_launch(lbmx_lattice *lat, int mode);
/* Steps per launch of a long run of this lattice under its mode (1 when the
 * lattice takes one step per launch). */
LBMX_API int lbmx_launch_depth(const lbmx_lattice *lat);
/* The same rule for a qualifying lattice of this size and precision, on the host
 * (no device needed): *steps_per_launch, and the output columns (*columns) and
 * rows (*rows) of one workgroup's strip segment (row streaming) or tile (the
 * tile form); both 0 for one step per launch. */
LBMX_API int lbmx_launch_plan(int Nx, int Ny, int precision, int mode, int *steps_per_launch,
                              int *columns, int *rows);

/* Time-domain impedance boundary (TDIBC) driven on the device every step, the
 * scheme of cases/nit/nit.c:13-47 and :275-303: before each step the wall
 * pressure p' = cs*cs*(rho - rho0)*cp is taken from the lagged density of the
 * probe line, an explicit-Euler auxiliary ODE (N=2 real poles, N=2 complex pole
 * pairs, updated in place as nit.c:31 does) gives the wall velocity
 * v = Zinf p' - A.phi - 2(B.psi + C.chi), and value = gain * (v / cu) is written
 * into the driven boundary's val1 or val2 at every active edge node (0 elsewhere).
 * Lumped mode (per_node = 0): one ODE on the probe line's mean density (nit).
 * Per-node mode: one ODE per edge node on the density of its probe-line node.
 * All ODE state is FP64 on the device; the history records, per step, the
 * (t, p', v) row nit.c:273 prints (values of the previous step, as it prints them). */
typedef struct {
    int bc;                 /* boundary-table entry whose values are driven */
    int which;              /* 1: val1, 2: val2 */
    int probe_is_column;    /* probe line is column probe_index (else row probe_index) */
    int probe_index;
    int per_node;
    int record_node;        /* per-node mode: the edge node whose (p', v) is recorded */
    int capacity;           /* history rows kept (steps) */
    double rho0, cs, cp, dt, cu, gain;
    double Zinf, A[2], lambdas[2], B[2], C[2], alpha[2], beta[2];
    const int *active;      /* edge-length flags (NULL: all active) */
} lbmx_tdibc;

LBMX_API int lbmx_set_tdibc(lbmx_lattice *lat, const lbmx_tdibc *cfg);
/* Copies min(rows, steps recorded) history rows (t, p', v) into rec[rows][3];
 * returns the number of rows copied. */
LBMX_API int lbmx_tdibc_history(lbmx_lattice *lat, double *rec, int rows);

/* Density probes: every step records, on the device, the density at n nodes as
 * the reference's rho array holds it at the start of that step (the previous
 * step's compute_macroscopic_fields output).  capacity = steps kept. */
LBMX_API int lbmx_set_probes(lbmx_lattice *lat, const int *x, const int *y, int n, int capacity);
/* Copies min(rows, steps recorded) rows of n densities into rec[rows][n];
 * returns the number of rows copied. */
LBMX_API int lbmx_probe_history(lbmx_lattice *lat, double *rec, int rows);

/* lbmx_step(nsteps) bracketed by HIP events recorded on the lattice's stream;
 * *ms receives the device time between them (blocks until done). */
LBMX_API int lbmx_time_steps(lbmx_lattice *lat, int nsteps, double *ms);

/* Force record on the device: from the next step on, every step appends the row
 * {t = row*dt, L*cf, D*cf} -- the record the cylinder driver writes per step
 * (cases/cylinder/cylinder.c:162-185) -- without a host read per step.
 * capacity 0 turns it off. */
LBMX_API int lbmx_set_force_record(lbmx_lattice *lat, double dt, double cf, int capacity);
/* Copies up to capacity recorded rows (3 doubles each); returns the row count. */
LBMX_API int lbmx_force_record(lbmx_lattice *lat, double *rows, int capacity);

/* Momentum-exchange force of the last step, as main_lbm returns it through
 * (L, D) (main_lbm.c:43): D = x-component, L = y-component, summed over every
 * fluid->solid link, walls included (compute_force_mem.c:26-58). */
LBMX_API int lbmx_forces(lbmx_lattice *lat, double *L, double *D);

/* Force on one body (an extension: the reference has only the total above, which
 * includes the channel walls).  body is Ny x Nx, nonzero on the solid nodes of the
 * body, or NULL for none; links into those nodes are summed separately from the
 * next step on, with no change to the total.  lbmx_set_geometry clears it.
 * lbmx_body_forces returns the last step's body force like lbmx_forces;
 * lbmx_body_force_record the rows {t, L_body*cf, D_body*cf} recorded alongside
 * lbmx_force_record's. */
LBMX_API int lbmx_set_force_body(lbmx_lattice *lat, const int *body);
LBMX_API int lbmx_body_forces(lbmx_lattice *lat, double *L, double *D);
LBMX_API int lbmx_body_force_record(lbmx_lattice *lat, double *rows, int capacity);

/* Moments of the current state exactly as compute_macroscopic_fields leaves
 * them after the last step (NaN on solid nodes).  Any pointer may be NULL. */
LBMX_API int lbmx_moments(lbmx_lattice *lat, double *rho, double *u, double *v);

/* FP64 totals of the moments lbmx_moments would return, over the fluid nodes:
 * mass = sum rho, energy = sum rho (u^2 + v^2) / 2, in a fixed summation order
 * (rows dealt to at most 1,024 blocks, a pairwise tree in each, then over the
 * blocks), so the result depends only on the field -- the full-grid reductions
 * the reference's drivers form on the host (cases/poiseuille/poiseuille.c:126-137),
 * here to state how far a reduced-precision run drifts from an FP64 one. */
LBMX_API int lbmx_totals(lbmx_lattice *lat, double *mass, double *energy);

/* No counterpart in the reference.  lbmx_totals for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops): the same
 * kernels and summation order over caller-owned device arrays, so the totals of
 * a field equal lbmx_totals' bit for bit.  rho, u, v: device arrays [Ny][Nx]
 * (row-major, the reference's layout); mask: device int [Ny][Nx], nonzero =
 * solid (skipped), or NULL for an all-fluid lattice; out: device memory for two
 * doubles {mass, energy}.  Enqueued on `stream` (a hipStream_t; NULL = the null
 * stream); only kernel launches, no allocation, no copy, no synchronization.
 * Its scratch is a fixed LBMX_NOALLOC_SCRATCH_BYTES device array per GPU, held
 * until the process ends; it covers any lattice of 1 .. LBMX_TOTALS_NOALLOC_MAX_NODES
 * nodes (other sizes: LBMX_ERR_ARG).  Calls that overlap on one GPU (several
 * streams, or threads) share that array and race. */
#define LBMX_NOALLOC_SCRATCH_BYTES 16384
#define LBMX_TOTALS_NOALLOC_MAX_NODES 2147483647LL
LBMX_API int lbmx_totals_noalloc(int Nx, int Ny, const double *rho, const double *u,
                                 const double *v, const int *mask, double *out, void *stream);

/* Current populations: f receives what the driver's `f` points to after the
 * last step and swap (the newest state); f_new (may be NULL) what its `f_new`
 * points to (the post-collision field of the last step, stale entries included).
 * Both are the reference's exact buffer contents; without LBMX_EXACT_STATE the
 * second one is completed on the device by this call. */
LBMX_API int lbmx_download(lbmx_lattice *lat, double *f, double *f_new);

/* write_vtk_binary_2D (lbm_utils.h; write_vtk_binary_2D.c:47-78) of the lattice's
 * current rho/u/v -- the moments lbmx_moments would return -- with the velocity
 * scaled by cu and the density by crho.  The big-endian float32 blocks are packed
 * on the device, so only 16 B per node cross PCIe; the file is byte-identical to
 * the host writer's. */
LBMX_API int lbmx_write_vtk(lbmx_lattice *lat, const char *filename, double dx, double cu,
                            double crho);

/* Blocks until the lattice's queued work is done. */
LBMX_API int lbmx_synchronize(lbmx_lattice *lat);

/* Steps run since lbmx_upload, and device bytes held by the lattice. */
LBMX_API int64_t lbmx_steps_done(const lbmx_lattice *lat);
LBMX_API int64_t lbmx_device_bytes(const lbmx_lattice *lat);

/* How lbmx_step groups the work after streaming into launches (no counterpart in
 * the reference; host only, no device needed).  For a lattice of this size,
 * periodicity, solid mask (Ny x Nx, or NULL) and boundary table, launch_of[b]
 * receives the launch (0, 1, ...) that applies table entry b, or -1 for a leading
 * periodic entry the step leaves out; *edge_launch the launch of the stale-edge
 * bookkeeping, or -1 on a fully periodic lattice.  Entries sharing a launch touch
 * disjoint nodes.  Returns the number of launches or a negative error code. */
LBMX_API int lbmx_tail_plan(int Nx, int Ny, int per_x, int per_y, const int *mask,
                            const int *types, const int *index, int nbc, int *launch_of,
                            int *edge_launch);

/* Measured ceiling of a two-buffer step (no counterpart in the reference): GB/s
 * (bytes read + written) of a plain device copy of gib GiB, 16-byte loads and
 * non-temporal stores, 5 timed launches after 2 untimed ones.  Allocates and frees
 * two buffers of gib GiB. */
LBMX_API int lbmx_copy_rate(double gib, double *gbs);

/* Mei consistent initialization on the device (mei_initialization.c:7-142):
 * same arguments and in/out contract as the reference function with the
 * lattice tables implied.  f/f_new are read as the starting buffers and
 * written back; rho/u/v receive the reference's outputs.  Returns the number
 * of iterations run (>= 1) or a negative error code. */
LBMX_API int lbmx_mei_initialization(int Nx, int Ny, double *f, double *f_new, const double *rho_in,
                            const double *U_in, const double *V_in, double *rho, double *u,
                            double *v, const double *omega, int per_x, int per_y,
                            int init_iter_max, double tol_rho, const int *solid_mask);

/* Single-stage entry points behind the reference's stand-alone functions
 * (lbm_core.h:33-92, lbm_bc.h:4-355).  Each uploads its host arrays, runs the
 * stage on the device with the reference's semantics and writes its outputs
 * back.  bc_type is an LBMX_BC_* code. */
LBMX_API int lbmx_stage_collision(int Nx, int Ny, double *f, const double *rho, const double *u,
                         const double *v, const double *omega, const int *mask, double Fx,
                         double Fy);
LBMX_API int lbmx_stage_streaming(int Nx, int Ny, const double *f, double *f_new, const int *mask,
                         int per_x, int per_y);
LBMX_API int lbmx_stage_bounce_back(int Nx, int Ny, const double *f, double *f_new, const int *mask,
                           int per_x, int per_y, int use_ibb, const double *phi);
LBMX_API int lbmx_stage_force(int Nx, int Ny, const double *f, const double *f_new, const int *mask,
                     int per_x, int per_y, double *Fx, double *Fy);
LBMX_API int lbmx_stage_moments(int Nx, int Ny, const double *f_new, const int *mask, double *rho,
                       double *u, double *v, double Fx, double Fy);
LBMX_API int lbmx_stage_bc(int Nx, int Ny, int bc_type, int index, const double *val1,
                  const double *val2, const double *f, double *f_new, const double *rho,
                  const double *u, const double *v, const int *mask);

/* Frees the device lattice the stand-alone stages cache between calls. */
LBMX_API void lbmx_stage_release(void);

/* Extensions of the drop-in API (no counterpart in the reference):
 * lbm_hip_sync writes the device state of the running main_lbm sequence back
 * into the caller's f/f_new (for callers that read populations mid-run);
 * lbm_hip_release ends that sequence and frees its device memory. */
LBMX_API void lbm_hip_sync(void);
LBMX_API void lbm_hip_release(void);

#ifdef __cplusplus
}
#endif
#endif
