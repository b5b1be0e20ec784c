/*
 * icls_oracle.h -- CPU restatement of the reference ICLS step (TEST INFRASTRUCTURE).
 *
 * This is the checker for the HIP library, not part of it: only tests/, the smoke
 * check and the CPU-baseline leg of bench.py load it.  The product library never
 * links or calls it.
 *
 * Arrays are stored exactly as the reference declares them (Fortran order, i fastest,
 * with explicit halos), and periodic halos are filled by explicit copies the way the
 * single-rank MPI exchanges of src/bound.f90:668-797 do.
 */
#ifndef ICLS_ORACLE_H
#define ICLS_ORACLE_H
#include "../include/icls.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_state oracle_state;

oracle_state* oracle_create(const icls_params* p);
void   oracle_destroy(oracle_state* S);
int    oracle_init_program(oracle_state* S);
int    oracle_init_derived(oracle_state* S);
int    oracle_step(oracle_state* S, int istep, int begin);

/* the operators of src/main.f90:103-291, one by one */
int    oracle_ls_main(oracle_state* S, int istep);
double oracle_average_stream_vel(oracle_state* S);
int    oracle_ab2(oracle_state* S);
int    oracle_fillps(oracle_state* S);
int    oracle_gfm4p(oracle_state* S);
int    oracle_add_pjump(oracle_state* S);
int    oracle_solver2d(oracle_state* S);
int    oracle_solver2d_cos(oracle_state* S, int bcy);   /* 1 Walls, 2 InOut, 3 OutOut */
void   oracle_vcosqb(int n, double* x);
void   oracle_vcosqf(int n, double* x);
int    oracle_boundp(oracle_state* S, int field);
int    oracle_correc(oracle_state* S);
int    oracle_bounduvw(oracle_state* S, int which);
int    oracle_update_pressure(oracle_state* S, int first_step);
int    oracle_average_vel_at_center(oracle_state* S);
int    oracle_chkdiv(oracle_state* S, int istep, double* divtot, double* divmax, int* ijk);
double oracle_chkdt(oracle_state* S);
void   oracle_ab2_coeffs(double dt, double dt_old, double c[4]);
double oracle_max_vel_at_center(oracle_state* S);

/* sub-steps, exposed for unit pins */
int    oracle_tube_construction(oracle_state* S, int l);
int    oracle_ls_advection(oracle_state* S, int l);
int    oracle_ls_mass_correction(oracle_state* S, int istep, int l);
int    oracle_inflation(oracle_state* S, int l);
int    oracle_ls_reinit(oracle_state* S, int l, int maxiter, int* iter, double* residual);
int    oracle_update_materials(oracle_state* S);
int    oracle_get_curvature(oracle_state* S, int l);       /* into curv_cmn + normal */
int    oracle_re_curvature(oracle_state* S, int l);        /* curv_cmn -> curv_cmn   */
int    oracle_mom_main(oracle_state* S, double* ru, double* rv, double* rw);
int    oracle_eval_adv(oracle_state* S, int l, int scheme, double* dfdt); /* interior, band only */
int    oracle_eval_reinit_rhs(oracle_state* S, int l, double* rhs);       /* LS shape */
int    oracle_band_count(oracle_state* S);
int    oracle_band_mask(oracle_state* S, int* mask);                        /* interior */
double oracle_ipot_dmin(oracle_state* S, int m, int n);

/* field access: returns a pointer into the oracle's own array */
double* oracle_field(oracle_state* S, int field, int l, int* shape, long* count);
double  oracle_scalar(oracle_state* S, int which);
void    oracle_set_scalar(oracle_state* S, int which, double v);
double  oracle_volume(oracle_state* S, int l, double* init_volume);
void    oracle_set_volume(oracle_state* S, int l, double volume, double init_volume);
const double* oracle_pseudo_inv_A(oracle_state* S);          /* 10 x 27, column-major */

/* stand-alone numerics (no state) */
void   oracle_rfft_forward(int n, double* r);    /* VRFFTF semantics, M = 1 */
void   oracle_rfft_backward(int n, double* r);   /* VRFFTB semantics, M = 1 */
void   oracle_pseudo_inverse(double dx, double dy, double dz, double* out /* 10x27 col-major */);
double oracle_heaviside_trig(double phi, double eps);
double oracle_dirac_trig(double phi, double eps);
double oracle_icurvature(int nb, const double* ph, const double* cv);
void   oracle_thomas(int kmax, double dzi, double lambda, double* p);

#ifdef __cplusplus
}
#endif
#endif
