/*
 * cnd.h -- C ABI of libcnd, the MI355X (gfx950) implementation of Canard's
 * compressible ILES time step.
 *
 * The reference has no FFI: its per-block step is inline Fortran in
 * src/mo_canard_driver.f90:226-431, calling type-bound procedures of
 * t_numerics / t_physics / t_grid / t_domdcomp.  Each entry point below
 * names the reference routine(s) it replaces.  A Fortran driver binds these
 * through iso_c_binding (canard_amd/fortran/mo_cnd.f90); Python binds them
 * through ctypes (canard_amd/_lib.py).
 *
 * Conventions (all as in the reference):
 *  - a block of lxi x let x lze cells stores (lxi+1)(let+1)(lze+1) points,
 *    linear index l = i + (lxi+1)*(j + (let+1)*k), xi fastest
 *    (src/mo_utils.f90:54-69, src/mo_domdcomp.f90:216);
 *  - fields are structure-of-arrays FP64: q(0:lmx,5) = (rho, rho u, rho v,
 *    rho w, rho E), xyz(0:lmx,3) (src/mo_canard_driver.f90:134-137);
 *  - boundary codes 10/20/25/30/35/40/45 (src/mo_parameters.f90:43-49);
 *    an interface face (30/35/40/45) names its partner block in mcd, and the
 *    partner's opposite face must name it back with the same code.
 *
 * Host pointers are borrowed for the duration of a call and copied.  Every
 * entry returns CND_OK or a negative error code; cnd_last_error() gives the
 * message.  Nothing aborts across the ABI.  One host thread per context;
 * work is ordered on the context's HIP stream.
 */
#ifndef CND_H
#define CND_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CND_OK = 0,
    CND_EINVAL = -1,   /* bad argument or unsupported topology      */
    CND_EDEVICE = -2,  /* HIP runtime error                         */
    CND_ESTATE = -3,   /* call out of order (e.g. step before grid) */
    CND_ENOMEM = -4
};

/* Precision of the step: CND_FP64; CND_MIXED, FP32 line solves with one
 * FP64 residual refinement and the stage's velocity and temperature
 * gradients stored in FP32 (state, metrics, fluxes, right-hand sides, RK
 * update and characteristic boundaries stay FP64).                         */
enum { CND_FP64 = 0, CND_MIXED = 1 };

typedef struct cnd_ctx cnd_ctx;

/* One block: mo_domdcomp.f90 t_domdcomp (lxi/let/lze, nbc, mcd) with one
 * rank per block (nbpc = 1,1,1).  inst groups blocks into independent
 * instances of an ensemble (each with its own parameters, time step and
 * clock); a plain multi-block case has every block in instance 0.          */
typedef struct {
    int lxi, let, lze;
    int nbc[3][2];
    int mcd[3][2];
    int inst;
} cnd_block;

/* Namelist parameters of one instance: nml_numerics (fltk, fltrbc),
 * nml_physics (reoo, tempoo, amach1-3, timf, nsmf), nml_canard (cfl),
 * nml_sponge / nml_gridgen (szco, szth0, szth1, doml0, doml1), and the
 * VISCOUS build switch (src/CMakeLists.txt:26-28).                          */
typedef struct {
    double fltk, fltrbc;
    double reoo, tempoo, amach1, amach2, amach3, timf;
    int    nsmf;
    double cfl;
    int    viscous;
    int    sponge;
    double szco, szth0, szth1, doml0, doml1;
} cnd_params;

/* Replaces the set-up half of canard_driver (mo_canard_driver.f90:103-148):
 * domain tables, numerics coefficients (init_extracoeff_bounds, init_penta,
 * mo_numerics.f90:114-249) and device allocation.  precision = CND_FP64 or
 * CND_MIXED (FP32 sweeps with one FP64 refinement pass).                   */
int  cnd_create(const cnd_params *params, int ninst, const cnd_block *blocks,
                int nblocks, int precision, cnd_ctx **ctx);
void cnd_destroy(cnd_ctx *ctx);
const char *cnd_last_error(const cnd_ctx *ctx);   /* ctx may be NULL */

/* Grid of block b, xyz(0:lmx,3).  Replaces read_grid_parallel
 * (mo_io.f90:151-180).                                                     */
int cnd_set_grid(cnd_ctx *ctx, int b, const double *xyz);
/* After every block's grid is set: grid metrics on the device
 * (calc_grid_metrics, mo_grid.f90:52-189), GCBC weights (gcbc_init,
 * mo_gcbc.f90:40-131), sponge lists (spongeup, mo_sponge.f90:34-71) and
 * junction tables (search_point/search_line, mo_domdcomp.f90:269-437).     */
int cnd_init_metrics(cnd_ctx *ctx);
/* Grid of block b generated on the device instead of uploaded
 * (mo_gridgen.f90:510-654, makegrid's gridtoy and gridtoy_multi), or the
 * sinusoidally sheared Cartesian block of the multi-block benchmark:
 *  CND_GRID_TOY:       par = lxi0, let0, lze0, doml0, doml1, domh, span;
 *      x = -doml0 + i (doml1 + doml0)/lxi0, y = -domh + j 2 domh/let0,
 *      z = span ((lze0 - k)/lze0 - 1/2)
 *  CND_GRID_TOY_MULTI: par = lxi0, lxi1, lxi2, let0, lze0, doml0, doml1,
 *      domh, span, mb (block mb = 0..5 of the 3 x 2 arrangement)
 *  CND_GRID_SHEARED:   par = x0, h, amp, lx, ly;
 *      x = x0 + i h + amp lx sin(2 pi j h / ly), y = j h, z = k h
 * cnd_get_grid copies block b's coordinates back, xyz(0:lmx,3).           */
enum { CND_GRID_TOY = 0, CND_GRID_TOY_MULTI = 1, CND_GRID_SHEARED = 2 };
int cnd_make_grid(cnd_ctx *ctx, int b, int kind, const double *par, int npar);
int cnd_get_grid(cnd_ctx *ctx, int b, double *xyz);
int cnd_get_metrics(cnd_ctx *ctx, int b, double *xim, double *etm,
                    double *zem, double *yaco);

/* State q(0:lmx,5) of block b (initialo / read_restart_file,
 * mo_physics.f90:117-140, mo_io.f90:184-229).                              */
int cnd_set_state(cnd_ctx *ctx, int b, const double *q);
int cnd_get_state(cnd_ctx *ctx, int b, double *q);    /* synchronizes */

/* Time-loop clock of instance inst: n, ndt, dt, dts, dte, timo
 * (mo_canard_driver.f90:204-210; restart record 1-2, mo_io.f90:209-216).   */
int cnd_set_clock(cnd_ctx *ctx, int inst, int n, int ndt, double dt,
                  double dts, double dte, double timo);
int cnd_get_clock(cnd_ctx *ctx, int inst, int *n, int *ndt, double *dt,
                  double *dts, double *dte, double *timo);  /* synchronizes */

/* nsteps whole time steps of every instance (mo_canard_driver.f90:226-371:
 * filter phase, four RK stages with time step, viscous stresses, fluxes,
 * GCBC, sponge, update, wall condition and junction/surface averaging,
 * then n = n+1, timo = timo+dt).  Asynchronous; dt stays on the device.    */
int cnd_step(cnd_ctx *ctx, int nsteps);
/* The two halves of one step, for callers that interleave their own work:
 * the filter phase with qo = qa (mo_canard_driver.f90:233-252) and stage
 * nk = 1..4 (mo_canard_driver.f90:257-363).  cnd_end_step advances n and
 * timo (mo_canard_driver.f90:369-370).                                     */
int cnd_filter(cnd_ctx *ctx);
int cnd_rk_stage(cnd_ctx *ctx, int nk);
int cnd_end_step(cnd_ctx *ctx);

/* Capture each kind of whole step in a hipGraph and replay it (1) or launch
 * kernel by kernel (0).                                                    */
int cnd_use_graph(cnd_ctx *ctx, int enable);
int cnd_synchronize(cnd_ctx *ctx);

/* Single sweeps over every block, for tests and callers of the numerics
 * alone.  in/out/f are arrays of nblocks host pointers to (0:lmx) fields.
 * cnd_deriv_sweep = mpigo (nt=0) + deriv along dir (0=xi,1=eta,2=zeta)
 * (mo_numerics.f90:429-565, 709-886); cnd_filter_sweep = mpigo (nt=1) +
 * filte, adding the filter correction in place (mo_numerics.f90:890-987). */
int cnd_deriv_sweep(cnd_ctx *ctx, int dir, const double *const *in,
                    double *const *out);
int cnd_filter_sweep(cnd_ctx *ctx, int dir, double *const *f);

/* Section timers, the device counterpart of the reference's timer table
 * (mo_timer.f90:84-98): HIP events on the context stream bracket every
 * kernel group while profiling is on (not inside graph replays).
 * Sections: CND_SEC_FILTER (filter sweeps), CND_SEC_TSTEP (time step),
 * CND_SEC_GRAD (velocity/temperature gradient sweeps), CND_SEC_STRESS
 * (stress assembly), CND_SEC_FLUX (flux sweeps), CND_SEC_GCBC,
 * CND_SEC_UPDATE (sponge + RK update), CND_SEC_AVG (wall, junction and
 * surface averaging).  cnd_profile_read synchronizes and returns the summed
 * milliseconds and kernel launches of one section since the last enable.   */
enum { CND_SEC_FILTER = 0, CND_SEC_TSTEP, CND_SEC_GRAD, CND_SEC_STRESS, CND_SEC_FLUX,
       CND_SEC_GCBC, CND_SEC_UPDATE, CND_SEC_AVG, CND_NSEC };
int cnd_profile(cnd_ctx *ctx, int enable);
int cnd_profile_read(cnd_ctx *ctx, int section, double *ms, long long *launches);

/* Recording of intermediate results (mo_canard_driver.f90:373-425).
 * cnd_record(ctx, 1) starts accumulating, at the end of every step, the
 * trapezoid sum qb += dt/2 (qo + qa) of the filtered and the new state and
 * each instance's recorded time dtsum += dt (0 stops; the first enable
 * allocates 5 fields and clears them).  cnd_record_read writes the output
 * variables of block b as 5 x npts FP32 values, variable after variable in
 * the state's point order: density, the three velocity components plus
 * the frame velocity, and pressure, of the time average qb / dtsum -- of
 * the filtered state qo when the instance has taken exactly one step, as
 * the reference's first output.  cnd_record_reset clears qb and dtsum, as
 * the reference does after writing an output.  The trigger (which steps
 * write an output, `ndata`/`tmax`) stays with the caller.                  */
int cnd_record(cnd_ctx *ctx, int on);
int cnd_record_read(cnd_ctx *ctx, int block, float *out);
int cnd_record_reset(cnd_ctx *ctx);

/* Output of results (mo_canard_driver.f90:173-195, 286-291, 373-425;
 * write_output_file / techead, mo_io.f90:345-529).
 * cnd_output_schedule sets instance inst's output times (nml_canard tmax,
 * ndata): output k (k = 0, 1, ...) is due on the step whose interval
 * [timo, timo + dt] holds k tmax / ndata, decided on the device at the
 * step's first stage; ndata < 1 turns it off (the default).  The count
 * restarts at -1, as the reference's ndati does at start-up.
 * cnd_output_due tells whether the step just taken was such a step (*due
 * 1 or 0), the output's number (the reference's ndati) and its time
 * timo - dtsum/2; synchronizes.
 * cnd_output_write writes the Tecplot file of instance inst's recorded
 * output (one zone per block, variables rho, u, v, w, p in FP32 as
 * cnd_record_read forms them, title nNNN) and then clears the instance's
 * time sums, as the reference does after an output.  async = 0 writes
 * before returning; async = 1 returns at once: the variables are formed on
 * the device in stream order, copied to pinned host memory on a second
 * stream and written by a host thread while later steps run (the I/O
 * server ranks of mo_io_server.f90:224-261); cnd_output_wait joins the
 * write and returns its status.  Any later output call waits for it too.
 * cnd_output_grid writes the grid file (file type 1, title "grid",
 * variables x, y, z) of instance inst.                                     */
int cnd_output_schedule(cnd_ctx *ctx, int inst, double tmax, int ndata);
int cnd_output_due(cnd_ctx *ctx, int inst, int *due, int *ndati, double *time);
int cnd_output_write(cnd_ctx *ctx, int inst, const char *path, int async);
int cnd_output_grid(cnd_ctx *ctx, int inst, const char *path);
int cnd_output_wait(cnd_ctx *ctx);
/* Tecplot binary (#!TDV112) file with nzones ordered zones, host data:
 * file_type 1 (grid) or 2 (solution), nvar variable names, dims[3 z + d]
 * points of zone z along d, data[z] = nvar x points FP32 values, variable
 * after variable, i fastest; zones are named z000, z001, ... with strand
 * z + 1 and the given solution time, and carry each variable's minimum and
 * maximum (techead + vminmax, mo_io.f90:417-562).  No device work.         */
int cnd_write_plt(const char *path, const char *title, int file_type, int nvar,
                  const char *const *names, int nzones, const int *dims,
                  double time, const float *const *data);

/* Sweep-shape autotuner (no counterpart in the reference, which has no
 * device kernels).  Kinds: CND_KIND_FILTER (filter phase), CND_KIND_GRAD
 * (gradient sweeps), CND_KIND_FLUX (flux sweeps), CND_KIND_FIELD (single
 * field sweeps: metrics, cnd_deriv_sweep).  Before the first step (or on
 * cnd_tune) every compiled shape that fits the blocks -- segments of L
 * points (lanes per pencil), pencils per workgroup, staging run length --
 * and, in FP64, the tiled one-lane-per-pencil sweep are timed on the
 * context's arrays per kind and direction, and the fastest is kept; all of
 * them give the same values up to rounding.  CANARD_TUNE=0 in the
 * environment keeps the defaults.  cnd_tuning reports the choice: shape
 * index (-1: tiled), segment length L (0: tiled), pencils per workgroup,
 * segments per pencil in the launch, and the measured milliseconds (0 if
 * not tuned).                                                              */
enum { CND_KIND_FILTER = 0, CND_KIND_GRAD, CND_KIND_FLUX, CND_KIND_FIELD, CND_NKIND };
int cnd_tune(cnd_ctx *ctx);
/* Pin the shape of one kind and direction (-1: tiled; FP64 only); turns
 * the automatic tuning off for this context.                               */
int cnd_set_shape(cnd_ctx *ctx, int kind, int dir, int shape);
/* Compiled shape number shape (0, 1, ...): segment length, pencils per
 * workgroup, staging run length, and whether it is a xi (1) or eta/zeta (0)
 * shape; CND_EINVAL past the last one.                                     */
int cnd_shape_info(int shape, int *seg_len, int *pencils, int *run, int *xi);
int cnd_tuning(const cnd_ctx *ctx, int kind, int dir, int *shape, int *seg_len,
               int *pencils, int *segments, double *ms);

/* Learned closure (no counterpart in the reference): a three-layer
 * 3x3x3 convolution network, 5 -> 32 -> 32 -> 5 channels with ReLU, on a
 * periodic n^3 coarse grid, evaluated and trained on the BF16 matrix cores
 * with FP32 accumulation.  It maps the (normalised) coarse state after a
 * step to a correction towards the box-filtered fine solution; the compact
 * filter stays in every step (a case without a network is the plain
 * scheme).  Parameters: cnd_net_param_count() FP32 values, W1[32][27][5],
 * b1[32], W2[32][27][32], b2[32], W3[5][27][32], b3[5], tap t =
 * (dx+1) + 3 (dy+1) + 9 (dz+1).  Coarse blocks have n+1 points per
 * direction (periodic duplicates).                                         */
typedef struct cnd_net cnd_net;
int  cnd_net_param_count(void);
/* n: coarse cells per direction (multiple of 8); capacity: training pairs
 * kept on the device; seed: initial weights.                               */
int  cnd_net_create(int n, int capacity, unsigned long long seed, cnd_net **net);
void cnd_net_destroy(cnd_net *net);
const char *cnd_net_error(const cnd_net *net);   /* net may be NULL */
int  cnd_net_get_params(cnd_net *net, float *w);
int  cnd_net_set_params(cnd_net *net, const float *w);
/* Input x = (q - in_mean) / in_scale; correction = y * out_scale (5 each). */
int  cnd_net_set_norm(cnd_net *net, const double *in_mean, const double *in_scale, const double *out_scale);
/* Store a training pair: input = state of block bc of coarse, target =
 * state of block bt of target minus it.  Returns the pair's index (>= 0).  */
int  cnd_net_add_pair(cnd_net *net, cnd_ctx *coarse, int bc, cnd_ctx *target, int bt);
/* The same from host arrays in normalised units, [n^3][5] each.           */
int  cnd_net_add_pair_host(cnd_net *net, const float *x, const float *target);
int  cnd_net_pairs(const cnd_net *net);
/* iters Adam steps (learning rate lr) on one stored pair each, drawn with
 * seed; mean_loss = mean square error over those steps.                    */
int  cnd_net_train(cnd_net *net, int iters, double lr, unsigned long long seed, double *mean_loss);
int  cnd_net_eval(cnd_net *net, int first, int count, double *mean_loss);
/* q += correction(q) on block b of ctx (asynchronous on ctx's stream).     */
int  cnd_net_apply(cnd_net *net, cnd_ctx *ctx, int b);
/* The same as cnd_net_apply, allocating nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops).  The input, hidden activations and
 * output live in a fixed device array of the library (41 MiB per GPU, held
 * until the process ends) instead of the network's own buffers, so n must
 * be at most 64 (CND_EINVAL otherwise).  Every call on one GPU shares that
 * array: calls that overlap on one GPU (different networks, or one network
 * on unordered streams) race.                                             */
int  cnd_net_apply_noalloc(cnd_net *net, cnd_ctx *ctx, int b);
/* Host test hooks, normalised units, channel-last [n^3][5]: forward, and
 * the gradients of L = sum(gy * y) with respect to x and the parameters.   */
int  cnd_net_forward_host(cnd_net *net, const float *x, float *y);
int  cnd_net_grad_host(cnd_net *net, const float *x, const float *gy, float *gx, float *gparams);
/* State of coarse block bc = top-hat filter (width r = nf / n fine cells,
 * trapezoidal end weights) of periodic fine block bf sampled at the coarse
 * points; cubic blocks, r even.                                            */
int  cnd_box_filter(cnd_ctx *fine, int bf, cnd_ctx *coarse, int bc);

/* Device stream the context works on (a hipStream_t), for callers that
 * order their own work against it.                                         */
void *cnd_stream(cnd_ctx *ctx);
/* Number of stored points summed over blocks, and nominal cells.          */
long long cnd_points(const cnd_ctx *ctx);
long long cnd_cells(const cnd_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
