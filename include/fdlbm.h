/*
 * fdlbm.h - C ABI of libfdlbm.so, the MI355X (gfx950) two-phase D2Q9
 * lattice-Boltzmann step of ebinan92/Fingering_dynamics.
 *
 * The reference has no FFI: its boundary is the Python object surface that
 * main() drives (lattice_boltzmann/fingering_periodic.py:369-479).  Each
 * entry point below replaces one piece of that surface; the reference
 * file:line it stands in for is given beside it.  Plain pointers and sizes
 * only; every host buffer is caller-owned, C-contiguous and copied.
 *
 * Errors: every int-returning call returns FD_OK (0) on success and a
 * negative FD_E* code otherwise; fd_last_error() then holds a message for
 * the calling thread.  No C++ exception crosses this boundary.
 * Threading: one handle per host thread; calls on a handle are serialised
 * on the handle's HIP stream.
 *
 * Array conventions (the reference's): row-major [y][x], y in [0, H) with
 * "north" = y+1, x in [0, W); populations are [9][H][W] in the D2Q9 order
 * e0=(0,0) e1=(1,0) e2=(0,1) e3=(-1,0) e4=(0,-1) e5=(1,1) e6=(-1,1)
 * e7=(-1,-1) e8=(1,-1) (fingering_periodic.py:50-89).
 */
#ifndef FDLBM_H
#define FDLBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FD_OK 0
#define FD_E_ARG -1      /* invalid argument (shape, null pointer, range)      */
#define FD_E_HIP -2      /* a HIP runtime call failed                           */
#define FD_E_STATE -3    /* call not valid in the handle's current state        */
#define FD_E_NODEV -4    /* no usable gfx950 device                             */

/* y boundary */
#define FD_Y_PERIODIC 0  /* np.roll wrap, fingering_periodic.py:219, 327-343   */
#define FD_Y_WALLS 1     /* half-way walls + psi_wall ghost rows, fingering.py  */
/* x boundary */
#define FD_X_ZOUHE 0     /* Zou-He inlet x=0 / outlet x=W-1, ghost columns +-1  */
#define FD_X_PERIODIC 1  /* build-original: periodic x                          */
/* inlet/outlet velocity profile */
#define FD_PROFILE_GAUSS 0 /* u0*exp(-t^2/2), fingering_periodic.py:270-271     */
#define FD_PROFILE_CONST 1 /* u0, fingering.py:299                              */
/* outlet kind */
#define FD_OUTLET_VELOCITY 0 /* fingering_periodic.py:305-324                   */
#define FD_OUTLET_PRESSURE 1 /* build-original: prescribed rho_out              */
/* initial-state quirk */
#define FD_INIT_MU0 0    /* mu = 0 at init, fingering_periodic.py:111-121       */
#define FD_INIT_WALLED 1 /* mu and uy computed at init, fingering.py:121-127    */

/* Node flag bits (one uint16 per node, fd_create's `links` argument uses the
 * low 8): bit (i-1) set => direction i (1..8) is overwritten at this node by
 * the same node's post-collision opposite population (half-way bounce-back,
 * bounce_back.py:25-167). */
#define FD_LINK_BITS 0x00FF

/* Lattice constants and boundary selection.  Module globals of
 * fingering_periodic.py:15-44 that Compute reads at call time. */
typedef struct fd_params {
    int32_t H, W;          /* rows, columns                                     */
    double tau;            /* g relaxation time (fingering_periodic.py:26)      */
    double eta_n;          /* lattice viscosity of the psi=+1 fluid (:34)       */
    double M;              /* viscosity ratio eta(psi=-1)/eta(psi=+1) (:22)     */
    double a, kappa;       /* free-energy constants (:38-39)                    */
    double gamma;          /* mobility (:40)                                    */
    double psi_wall;       /* wall order parameter (:18)                        */
    double u0;             /* inlet velocity amplitude (:36)                    */
    double rho_out;        /* pressure-outlet density (FD_OUTLET_PRESSURE)      */
    double outlet_f3_coef; /* 2/3 (fingering_periodic.py:317) or 1.5 (fingering.py:378) */
    double gx, gy;         /* build-original body force added to mu*grad psi    */
    int32_t y_bc;          /* FD_Y_*                                            */
    int32_t x_bc;          /* FD_X_*                                            */
    int32_t profile;       /* FD_PROFILE_*                                      */
    int32_t outlet;        /* FD_OUTLET_*                                       */
    int32_t init_mode;     /* FD_INIT_*                                         */
    int32_t wall_lo;       /* FD_Y_WALLS: row of the bottom wall (0, or 1 as fingering.py:573 ships) */
    int32_t wall_hi;       /* FD_Y_WALLS: row of the top wall (H-1, or H-2)     */
    double psi_visc_bound; /* > 0: |psi| <= b in the tau_mix interpolation    */
                           /* (build-original), 0: fingering_periodic.py:201-208 */
} fd_params;

typedef struct fd_handle fd_handle;

/* Create a solver for one H x W domain on the current HIP device.
 * Replaces Compute(mask) construction, fingering_periodic.py:48-121, together
 * with the mask set-up of :449-451 and the bounce-back mask lists of
 * bounce_back.py:89-167 / :25-86 (packed per node).
 *   solid  uint8[H*W], non-zero on solid nodes (block_psi_all == 1)
 *   links  uint8[H*W], bit (i-1) = direction i bounced at this node
 * The state is not initialised until fd_init. */
int fd_create(const fd_params* p, const uint8_t* solid, const uint8_t* links, fd_handle** out);

/* Initial state, fingering_periodic.py:90-121 (or fingering.py:95-127 with
 * FD_INIT_WALLED).  psi0 double[H*W] (values on solid nodes are replaced by
 * psi_wall); rho0 double[H*W] or NULL for rho = 1 on fluid nodes.  Resets
 * the step counter to 0. */
int fd_init(fd_handle* h, const double* psi0, const double* rho0);

/* Restart from given fields, double[H*W] each: the equilibrium state of
 * psi, rho and velocity (ux, uy) with mu and the force mu grad psi taken
 * from psi, collided once - the state a running flow had, with a new phase
 * field.  Resets the step counter to 0; reads before the first fd_step
 * return the given fields (and their equilibria as populations).  No
 * reference counterpart: used where a learned surrogate hands the phase
 * field back to the solver. */
int fd_init_fields(fd_handle* h, const double* psi, const double* rho, const double* ux, const double* uy);

/* Advance nsteps iterations of the main loop body
 * (fingering_periodic.py:454-479: collide, stream, bounce-back, walls,
 * Zou-He, moments).  Enqueued on the handle's stream; returns without
 * waiting. */
int fd_step(fd_handle* h, int32_t nsteps);

/* 1: fd_step replays long step loops from a captured HIP graph of two
 * launches; 0 (default): every launch is issued eagerly.  Identical results.
 * Measured on MI355X, replay is 1-2 us per iteration slower than eager
 * launches on a 400 x 100 channel and the same on large domains.  No
 * reference counterpart (a launch control of this implementation). */
int fd_use_graph(fd_handle* h, int32_t on);

/* Number of iterations done since fd_init. */
int64_t fd_steps_done(const fd_handle* h);

/* Macroscopic fields after the last completed iteration, float[H*W] each;
 * any pointer may be NULL.  Replaces the attributes cm.psi, cm.rho, cm.ux,
 * cm.uy read by fingering_periodic.py:486-487 (the reference keeps rho/ux/uy
 * compressed over fluid nodes; here solid nodes hold 0 and psi holds
 * psi_wall there).  Synchronises the handle's stream. */
int fd_get_fields(fd_handle* h, float* psi, float* rho, float* ux, float* uy);

/* The same fields restricted to columns [x0, x1), float[H*(x1-x0)] each
 * (row-major); any pointer may be NULL.  Only the 64-column strips holding
 * the window are evaluated, so reading the outlet column to detect
 * breakthrough (any fluid node of column W-1 with psi > 0) costs a small
 * fraction of an iteration.  Partial read of cm.psi / cm.ux ... of
 * fingering_periodic.py:486-487; no separate reference counterpart.
 * Synchronises the handle's stream. */
int fd_get_fields_window(fd_handle* h, int32_t x0, int32_t x1, float* psi, float* rho, float* ux, float* uy);

/* Post-stream populations f, g float[9*H*W] after the last completed
 * iteration (cm.f, cm.g).  Either pointer may be NULL.  Synchronises. */
int fd_get_populations(fd_handle* h, float* f, float* g);

/* Replace the state by given post-stream populations f, g float[9*H*W] of
 * the current iteration (the quantity fd_get_populations returns: after
 * streaming, bounce-back, walls and Zou-He).  Replaces the in-place writes
 * of cm.f / cm.g by the reference's loop (fingering_periodic.py:459-460,
 * 464-467).  The iteration count is kept (0 on a handle not yet
 * initialised); the next iteration collides exactly these populations, with
 * psi and its stencil taken from them and the Zou-He side values P of this
 * psi.  Reads before the next fd_step return these populations and their
 * fields.  With inert solid nodes (fd_solid_inert) the values given on solid
 * nodes are dropped (read back as 0).  Synchronises. */
int fd_set_populations(fd_handle* h, const float* f, const float* g);

/* FP64 sums over fluid nodes of the last completed iteration:
 * out[0] = sum rho, out[1] = sum psi, out[2] = number of fluid nodes whose
 * rho or psi is not finite.  No reference counterpart: the reference's only
 * health check is np.seterr(all='raise') (fingering_periodic.py:497).
 * Synchronises. */
int fd_phase_mass(fd_handle* h, double out[3]);

/* fd_phase_mass without any allocation: the FP64 partial sums go to a fixed
 * 4 MiB scratch array inside the library (131,072 slots of 32 B: one per
 * wave of the moments launch, or per 256 nodes before the first iteration),
 * one copy per GPU, held until the process ends.  No reference counterpart:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  A launch
 * shape (fd_set_tile) that needs more slots fails with FD_E_ARG.  Every call
 * on one GPU uses that GPU's array: calls that overlap on one GPU (other
 * host threads or handles) race.  Synchronises. */
int fd_phase_mass_noalloc(fd_handle* h, double out[3]);

/* Failure watchdog of the last completed iteration, FP64:
 * out[0] = number of fluid nodes whose rho or psi is not finite,
 * out[1] = min over fluid nodes of tau_mix - 1/2 (FP32 per node).  BGK needs
 * tau_mix > 1/2: out[1] <= 0 (or a huge value turning negative) means psi
 * has crossed the pole psi = -(1+M)/(M-1) of the viscosity interpolation
 * (fingering_periodic.py:201-208, :205); +inf if there is no fluid node.
 * No reference counterpart: the reference's only guard is
 * np.seterr(all='raise') (fingering_periodic.py:497).  Synchronises. */
int fd_watchdog(fd_handle* h, double out[2]);

/* Opaque resumable state (post-collision populations, Zou-He side buffer,
 * step counter; or, right after fd_set_populations, the given
 * populations).  fd_state_bytes gives the size; fd_save_state copies it
 * out, fd_load_state copies it back into a handle created with the same
 * fd_params and masks.  There is nothing to save between fd_init and the
 * first fd_step (FD_E_STATE): fd_init rebuilds that state.  No reference
 * counterpart (the reference keeps no checkpoints,
 * fingering_periodic.py:461-462). */
int64_t fd_state_bytes(const fd_handle* h);
int fd_save_state(fd_handle* h, void* buf, int64_t nbytes);
int fd_load_state(fd_handle* h, const void* buf, int64_t nbytes);

/* Device-array coupling (no reference counterpart: the reference keeps its
 * fields in numpy, fingering_periodic.py:472-479).  These take DEVICE
 * pointers, are enqueued on the handle's stream and return without
 * waiting; a caller that produces or consumes the arrays on the GPU runs
 * its own work on fd_stream(h) (a hipStream_t) to stay ordered.
 *   fd_fields_device       psi, rho, ux, uy of the last completed iteration
 *                          into float[H*W] device arrays (any may be NULL).
 *   fd_init_fields_device  fd_init_fields from float[H*W] device arrays
 *                          (all four required); the step count restarts at 0.
 *   fd_state_device_bytes / fd_save_state_device / fd_load_state_device
 *                          the resumable state (populations and Zou-He side
 *                          buffer) as a device blob; the step count is kept
 *                          by the caller and handed back on load (>= 1).
 *                          Saving needs an iteration since the last
 *                          fd_init* / fd_set_populations (FD_E_STATE). */
void* fd_stream(const fd_handle* h);
int fd_fields_device(fd_handle* h, float* psi, float* rho, float* ux, float* uy);
int fd_init_fields_device(fd_handle* h, const float* psi, const float* rho, const float* ux, const float* uy);
int64_t fd_state_device_bytes(const fd_handle* h);
int fd_save_state_device(fd_handle* h, void* dbuf, int64_t nbytes);
int fd_load_state_device(fd_handle* h, const void* dbuf, int64_t nbytes, int64_t steps_done);

/* Wait for the handle's queued work. */
int fd_synchronize(fd_handle* h);

/* Average device time per iteration (ms) of the last fd_step call, measured
 * with HIP events on the handle's stream; < 0 if unavailable. */
double fd_last_step_ms(const fd_handle* h);

/* Kernel launch shape: workgroups of waves_x waves (1, 2, 4 or 8), each
 * wave marching `rows` rows (>= 1) of its own 64-column strip.  The default
 * is 4 x 16, fewer rows (and 1 wave) on domains too small to fill the
 * device.  No reference counterpart (a launch-shape control of this
 * implementation). */
int fd_set_tile(fd_handle* h, int32_t waves_x, int32_t rows);

/* Iterations per kernel launch of fd_step: depth 2 advances pairs of
 * iterations with one launch that keeps the intermediate state on chip, so
 * the populations cross HBM once per two iterations; depth 3 does the same
 * for triples (two intermediate states on chip); depth 1 runs one launch
 * per iteration.  The results are identical bit for bit.  waves_x (1, 2 or
 * 4; depth 3 uses 1 or 2) and rows set the launch shape of that depth's
 * kernel (<= 0 keeps the current one).  Defaults: depth 3 at 1 wave x 16
 * rows on domains that fill the device at least twice over with such waves
 * (4096 x 2048 does), else depth 2 at 1 wave x 8 rows,
 * fewer rows on domains too small to fill the device.  Domains with H < 8,
 * and walled domains whose walls are not the edge rows (the rows 1 and H-2
 * of fingering.py:573), always run depth 1.  No reference counterpart (a
 * launch control of this implementation). */
int fd_set_fusion(fd_handle* h, int32_t depth, int32_t waves_x, int32_t rows);

/* The iterations per launch fd_step uses now (1 when the domain does not
 * allow more) and that kernel's launch shape; NULL outputs are skipped.
 * No reference counterpart. */
int fd_get_fusion(const fd_handle* h, int32_t* depth, int32_t* waves_x, int32_t* rows);

/* 1 if the handle runs with inert solid nodes, 0 if solid nodes stream
 * their populations on (store-and-forward, fingering_periodic.py:455-466),
 * -1 for NULL.  Inert needs a leak-free geometry: every solid -> fluid link
 * bounced (or rewritten by a boundary condition) and no solid node on a
 * Zou-He column.  Solid nodes are then never read and hold zeros, and
 * populations reported on solid nodes are 0.  In both modes a bounced link
 * costs no extra load (the pull reads the node's own opposite population).
 * No reference counterpart (a property of this implementation). */
int fd_solid_inert(const fd_handle* h);

/* HBM calibration: copies nbytes (>= 32 MiB, rounded down to 16 B) between
 * two device buffers the library allocates, `reps` times per launch shape,
 * with a streaming copy kernel (fd_copy_kernel: 16-byte non-temporal loads
 * and stores) at three grid sizes, and returns the best average milliseconds
 * per copy - the streaming rate this device reaches in practice.  Also the
 * yardstick that scales rocprofv3 FETCH_SIZE / WRITE_SIZE counters.  No
 * reference counterpart (a measurement aid). */
int fd_copy_probe(int64_t nbytes, int32_t reps, double* ms_per_copy);

/* Release every device buffer of the handle.  NULL is ignored.  A handle
 * that still belongs to an ensemble is kept (its buffers are in the
 * ensemble's device table): the call then does nothing and fd_last_error()
 * says so; destroy the ensemble first. */
void fd_destroy(fd_handle* h);

/* ---- Ensembles: many independent domains advanced together ------------
 * No reference counterpart: the reference runs one domain per process
 * (fingering_periodic.py:454-479).  An ensemble batches existing handles
 * (any sizes, parameters and boundary conditions, one device) so that one
 * kernel launch advances every domain by one iteration; the launches of a
 * step loop are replayed from a captured HIP graph.  Each handle keeps its
 * own state, and fd_get_fields / fd_phase_mass / fd_save_state ... on a
 * handle stay valid (they are ordered after the ensemble's work).  Destroy
 * the ensemble before its handles: fd_destroy refuses a member handle. */
typedef struct fd_ensemble fd_ensemble;

/* Batch n distinct, initialised handles of one device. */
int fd_ens_create(fd_handle* const* hs, int32_t n, fd_ensemble** out);

/* nsteps iterations of every domain: one launch per iteration covering all
 * of them (graph-replayed in pairs unless disabled).  All handles must be
 * at the same iteration count.  Asynchronous. */
int fd_ens_step(fd_ensemble* e, int32_t nsteps);

/* The same iterations as one launch per domain per iteration, in sequence
 * on the ensemble's stream: the baseline a batched launch is measured
 * against.  Results are identical to fd_ens_step. */
int fd_ens_step_sequential(fd_ensemble* e, int32_t nsteps);

/* Launch shape of the batched single-iteration kernel, as fd_set_tile
 * (default 4 x 16). */
int fd_ens_set_tile(fd_ensemble* e, int32_t waves_x, int32_t rows);

/* Iterations per batched launch, as fd_set_fusion: depth 3 (the default)
 * advances triples of iterations of every domain with one launch of the
 * three-iteration kernel, depth 2 pairs with the two-iteration kernel,
 * depth 1 one iteration per launch.  waves_x and rows set that depth's
 * launch shape (<= 0 keeps the current one; defaults 1 x 26 for depth 3,
 * 1 x 8 for depth 2; depth 3 takes 1 or 2 waves).  Groups run batched only
 * when every domain's geometry allows the fused kernels (fd_set_fusion);
 * otherwise every launch is of one iteration.  Results are identical bit
 * for bit in every case.
 * No reference counterpart (the reference runs one domain per process). */
int fd_ens_set_fusion(fd_ensemble* e, int32_t depth, int32_t waves_x, int32_t rows);

/* 1: replay step loops from a captured graph (default); 0: launch eagerly. */
int fd_ens_use_graph(fd_ensemble* e, int32_t on);

/* Wait for the ensemble's queued work. */
int fd_ens_synchronize(fd_ensemble* e);

/* Average device time per iteration (ms, all domains) of the last
 * fd_ens_step / fd_ens_step_sequential call; < 0 if unavailable. */
double fd_ens_last_step_ms(const fd_ensemble* e);

/* Release the ensemble (not its handles).  NULL is ignored. */
void fd_ens_destroy(fd_ensemble* e);

/* Message of the last failed call on this thread ("" if none). */
const char* fd_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* FDLBM_H */
