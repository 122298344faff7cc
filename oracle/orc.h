/*
 * orc.h -- CPU oracle for Canard's compressible ILES time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference's Fortran algorithm (src/mo_numerics.f90, src/mo_physics.f90,
 * src/mo_grid.f90, src/mo_gcbc.f90, src/mo_sponge.f90, src/mo_domdcomp.f90,
 * src/mo_canard_driver.f90).  It is used only by tests/, by the smoke check
 * and by the CPU-baseline leg of bench.py, always as the checker and never as
 * the thing measured or shipped.  The product (canard_amd, libcnd) never
 * links or calls it.
 *
 * Every block of the multi-block case lives in this one process; the MPI
 * traffic of the reference is emulated (face partial sums and face
 * exchanges by partner lookup, junction averaging by replaying the
 * reference's send/receive posting order with per-channel FIFO matching).
 */
#ifndef ORC_H
#define ORC_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double fltk;        /* filter cut-off as a fraction of pi (nml_numerics) */
    double fltrbc;      /* boundary filter parameter (nml_numerics)          */
    double reoo;        /* Reynolds number (nml_physics)                     */
    double tempoo;      /* free-stream temperature in K (nml_physics)        */
    double amach1, amach2, amach3;
    double timf;        /* moving-frame ramp time                            */
    int    nsmf;        /* 1: constant frame velocity, 0: ramp over timf    */
    double cfl;
    int    viscous;     /* 1: the VISCOUS build, 0: Euler                    */
    int    sponge;      /* 1: apply the sponge source                        */
    double szco, szth0, szth1, doml0, doml1;  /* sponge geometry            */
    int    quirk_send2; /* 1: keep the reference's persisting filter slot-2
                           update s = s + ((f - r) - s) (mo_numerics.f90:509);
                           0: store f - r directly                          */
} orc_params;

typedef struct {
    int lxi, let, lze;     /* cells per direction (points = cells + 1)      */
    int nbc[3][2];         /* boundary code per direction and side          */
    int mcd[3][2];         /* partner block per face, -1 if none            */
} orc_blockdesc;

typedef struct orc_case orc_case;

orc_case *orc_create(const orc_params *p, const orc_blockdesc *blk, int nblk);
void      orc_destroy(orc_case *c);
const char *orc_error(void);

int  orc_set_grid(orc_case *c, int b, const double *xyz);      /* (npt,3) SoA */
void orc_get_metrics(orc_case *c, int b, double *xim, double *etm,
                     double *zem, double *yaco);
void orc_set_state(orc_case *c, int b, const double *qa);      /* (npt,5) SoA */
void orc_get_state(orc_case *c, int b, double *qa);
void orc_set_clock(orc_case *c, int n, int ndt, double dt, double dts,
                   double dte, double timo);
void orc_get_clock(orc_case *c, int *n, int *ndt, double *dt, double *dts,
                   double *dte, double *timo);
int  orc_run(orc_case *c, int nsteps);          /* whole time steps        */
int  orc_filter_phase(orc_case *c);             /* filter + qo=qa          */
int  orc_stage(orc_case *c, int nk);            /* one RK stage, nk=1..4   */

/* Single sweeps over every block (fields given per block).  Partner partial
 * sums are formed from the same fields, as one mpigo call would.           */
void orc_deriv_all(orc_case *c, int nn, const double *const *in,
                   double *const *out);
void orc_filter_all(orc_case *c, int nn, double *const *f);

/* Coefficient tables for pinning tests.
 * out layout (doubles): alphf, betf, fa, fb, fc, albef[3][5] (ns=0 rows),
 * fbc[3][5], pbci[2][2][12], pbco[2][2][12], pbcot[2][2], sbci[4]       */
int  orc_coefs(orc_case *c, double *out, int nout);
/* LU factors of one block/direction/scheme: xl (n,2), xu (n,3) */
void orc_penta(orc_case *c, int b, int nn, int nt, double *xl, double *xu);
/* junction tables of one block: per corner partner count, per edge count  */
void orc_junctions(orc_case *c, int b, int *npoint8, int *nline12);

/* Test hooks: GCBC characteristic transforms at one state (row-major 5x5),
 * and the stage tail (wall update, junction and surface averaging) applied
 * to the current state with frame velocity umf.                            */
void orc_gcbc_mats(const double cm[3], const double q[5], const double umf[3], double *xtq2r, double *xtr2q);
void orc_stage_tail(orc_case *c, const double umf[3]);

void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
