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
