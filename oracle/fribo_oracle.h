/*
 * CPU oracle for the FRiBoIS2D fractional-step fluid step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Fortran (src/fluid_tsch.F90 and everything it calls on one MPI rank).  It is
 * used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg as the checker.  The product (fribois2d_amd, libfribo_gpu.so) never
 * links, loads or calls it.
 *
 * All arrays use the reference's Fortran layout (column-major, j fastest) and
 * extents on one rank (kstart=1, kend=n3m):
 *   q2,q3,pr,dph,q2f,q3f : (1:n2, -2:n3m+3)          mpi_routines.F90:428-441
 *   fey,fez              : (1:n2, -2:n3m+2)          utils_proc.F90:310-320
 *   ru2,ru3,rhs2,rhs3,rhs,qcap : (1:n2, 1:n3m)       mpi_routines.F90:450-466
 *   qb2s,qb3s,qb2n,qb3n,dqb2s,dqb3s,dqb2n,dqb3n,dq2x2o,dq3x2o : (1:n2)
 *   lmcc,lmvel : (nlm,2)   lmA : (nlm)
 *   lmind : (nlm,2,2)   eulins : (nlm,2,2,9)   phi,phipred : (nlm,2,9)   cl : (nlm,2)
 *   zmm, zmmt : (n3m, n3m) packed (leading dimension n3m)
 */
#ifndef FRIBO_ORACLE_H
#define FRIBO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n2, n3;          /* nodes in y (wall-normal) and z (streamwise) */
    int nsst;            /* 3 (RK3) or 1 (AB2)                          */
    int idtv;            /* 1: fixed dt (CFL checked), 0: CFL-controlled */
    int ibact;           /* immersed boundary active                   */
    int forst;           /* forcing repetitions                         */
    int wfun;            /* MLS weight function 1..4                    */
    int inslws, inslwn;  /* no-slip bottom / top wall                   */
    double alx3, rext2;  /* domain lengths z, y                         */
    double visc;         /* 1/Re                                        */
    double cou;          /* outflow Courant number                      */
    double q3uw, q3lw;   /* wall z-velocities                           */
    double dtmax, cflmax, cfllim;
    double radiusSD;     /* MLS support radius factor                   */
} orc_params;

typedef struct orc_state orc_state;

/* Error codes (match include/fribo_gpu.h). */
enum {
    ORC_OK = 0,
    ORC_ERR_DT_TOO_SMALL = 1,
    ORC_ERR_CFL_TOO_LARGE = 2,
    ORC_ERR_MARKER_OUTSIDE = 3,
    ORC_ERR_PARTITION_UNITY = 4,
    ORC_ERR_ARG = 5
};

orc_state *orc_create(const orc_params *p, const double *rc, const double *zz,
                      const double *zmm, const double *zmmt, const double *wr);
void orc_destroy(orc_state *s);

/* Element count of a named array (0 if unknown) and copies in/out. */
long orc_size(orc_state *s, const char *name);
int orc_get(orc_state *s, const char *name, void *out);
int orc_set(orc_state *s, const char *name, const void *in);

int orc_set_markers(orc_state *s, int nlm, const double *lmcc,
                    const double *lmvel, const double *lmA);

/* Index of the Neumann null mode of A_z found at creation (1-based), 0 if none. */
int orc_null_mode(orc_state *s);

/* fluid_tsch(ns): CFL/dt control and nsst substeps.  dt is in/out, dt_o in. */
int orc_fluid_tsch(orc_state *s, int ns, double *dt, double dt_o, double *cflm);

/* Finer-grained entry points used by the stage-by-stage parity tests. */
int orc_cfl_dt(orc_state *s, int ns, double *dt, double dt_o, double *cflm);
int orc_substep(orc_state *s, int n, int ns);

/* loads_transfer_c (loads_transfer.F90:4-162) after a step: pressure
 * press(nlm) and friction stress tau(nlm,2) at the markers from the probe
 * values; lmnor, lmacc (nlm,2) as rigid() leaves them, kprobe from
 * global_par.in.  checkbcs (ckbcs.F90:5-86): marker velocity errors and
 * their norms.  Both use the marker data of the last step. */
int orc_loads_transfer(orc_state *s, double kprobe, const double *lmnor, const double *lmacc, double *press,
                       double *tau);
int orc_checkbcs(orc_state *s, double *bcerrv, double *bcerrw, double *nbcv, double *nbcw);

/* The diagnostics ttime prints every step (ttime.F90:29-69,136): out[9] =
 * {vmax q2, vmax q3, vmin q2, vmin q3, max|div|, kinetic energy, vorticity
 * integral on z=18.5, cflm, q3avg}. */
int orc_diagnostics(orc_state *s, double *out);

/* Test hooks: DGTTRF+DGTTRS on one system, GEsolve on a 3x3 (column-major),
 * and phcalc_outflow applied to a (n2m x n3m) column-major right-hand side. */
int orc_test_dgtsv(int n, const double *dl, const double *d, const double *du, double *b);
int orc_test_gesolve(const double *a_colmajor, const double *b, double *x);
int orc_test_phcalc(orc_state *s, double *field);
double orc_test_weight1d(double rcv, int wfun);

/* Stage-level test hooks: set the step scalars {dt, beta, al, ga, ro}, and
 * run one named stage of the substep ("cfl", "boucdq", "boucqt", "explicit",
 * "lgrdistribution", "forcing1", "helsolve", "divg", "updvp", "prcalc", "pr_submean") on the
 * current state. */
int orc_set_scalars(orc_state *s, const double *v);
int orc_test_stage(orc_state *s, const char *name);

#ifdef __cplusplus
}
#endif
#endif
