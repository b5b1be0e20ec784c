/*
 * icls_oracle.c -- CPU restatement of the reference ICLS per-timestep step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library
 * (icls-release_amd/csrc); the library never links or calls it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load liboracle.
 *
 * Each function follows the reference routine cited in its comment, in the same
 * operation order (Fortran left-to-right evaluation, serial i-fastest reductions),
 * built with -ffp-contract=off so that no multiply-add is fused.  Arrays are stored
 * with the reference's explicit halos and lower bounds; the periodic halos are filled
 * by explicit copies, as the single-rank MPI_SENDRECV exchanges of
 * src/bound.f90:668-797 do.
 *
 * Deviations from the reference, all numerical-library substitutions:
 *   - VFFTPACK (libs/fft/fft.f) is replaced by a radix-2 / direct real DFT with the
 *     same half-complex packing and sqrt(1/N) scaling (fft.f:1128-1141, 1770);
 *     pinned against VFFTPACK compiled from source by oracle/Makefile.
 *   - LAPACK dgetrf/dgetri (misc.f90:566-573) for the 10x10 least-squares normal
 *     matrix is replaced by Gauss-Jordan in long double, rounded once to double;
 *     pinned against the exact rational pseudo-inverse in tests.
 */
#include "icls_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PI_F77 3.1415926535897932   /* PIMACHNEW, libs/fft/fft.f:1877-1889 */

/* The per-cell and per-line loops carry OMP_FOR.  Built with -fopenmp
 * (liboracle_omp.so) their iterations run on all host cores; every iteration writes
 * its own outputs, so the results are those of the serial build bit for bit.  The
 * reductions marked OMP_FOR_RED are integer sums, minima and maxima, which no order
 * changes; the floating-point sums stay serial, in their order. */
#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(static)")
#define OMP_PRAGMA_(x) _Pragma(#x)
#define OMP_FOR_RED(r) OMP_PRAGMA_(omp parallel for schedule(static) reduction(r))
#else
#define OMP_FOR
#define OMP_FOR_RED(r)
#endif

struct oracle_state {
    icls_params P;
    int imax, jmax, kmax, i1, j1, k1, lmax;
    int lsx, lsy, lsz, cx, cy, cz;
    long nls, nc, ni;
    double lx, ly, lz, dxi, dyi, dzi, dx, dy, dz;
    double NB2, rho0, visc, pi;
    double dt, dt_old, time, rho_avr, v_bulk, maxchk_glb, norm_old;
    double reinit_residual;
    /* C-halo arrays (0:i1,0:j1,0:k1) */
    double *unew, *vnew, *wnew, *pnew, *dudt, *dvdt, *dwdt;
    double *jiemian, *miu, *rho, *u_l, *v_l, *w_l, *slset, *curv_cmn;
    double *p_x, *p_y, *p_z, *p_xold, *p_yold, *p_zold;
    double *pold, *p_jump, *p;
    double **hnew_mls;
    /* interior arrays (1:imax,1:jmax,1:kmax) */
    double *uccnew, *vccnew, *wccnew, *rho_u, *rho_v, *rho_w;
    double *nrm_m, *nrm_x, *nrm_y, *nrm_z;
    double *duold, *dvold, *dwold;
    /* LS arrays (-2:i1+2,...) */
    double **lvset;
    /* narrow band */
    int nb_cnt, *nb_i, *nb_j, *nb_k;
    double *volume, *init_volume, *dmin;
    double pinvA[10 * 27];   /* pseudo_inv_A(10,27), column-major */
    double *a, *b, *c, *xyrt;
};

/* ---------------------------------------------------------------- indexing */
#define LS(i, j, k) ((long)((i) + 2) + (long)S->lsx * ((long)((j) + 2) + (long)S->lsy * (long)((k) + 2)))
#define CC(i, j, k) ((long)(i) + (long)S->cx * ((long)(j) + (long)S->cy * (long)(k)))
#define II(i, j, k) ((long)((i) - 1) + (long)S->imax * ((long)((j) - 1) + (long)S->jmax * (long)((k) - 1)))

static double* zalloc(long n) { return (double*)calloc((size_t)n, sizeof(double)); }

/* ---------------------------------------------------------------- halos (bound.f90) */
/* updthalos(var,1) and (var,2): bound.f90:668-716, datatypes initmpi.f90:144-147 */
static void updthalos(oracle_state* S, double* v) {
    for (int k = 0; k <= S->k1; k++)
        for (int j = 0; j <= S->j1; j++) {
            v[CC(S->i1, j, k)] = v[CC(1, j, k)];
            v[CC(0, j, k)] = v[CC(S->imax, j, k)];
        }
    for (int k = 0; k <= S->k1; k++)
        for (int i = 0; i <= S->i1; i++) {
            v[CC(i, S->j1, k)] = v[CC(i, 1, k)];
            v[CC(i, 0, k)] = v[CC(i, S->jmax, k)];
        }
}

/* updthalos_ls(var,1) and (var,2): bound.f90:760-797, datatypes initmpi.f90:156-159 */
static void updthalos_ls(oracle_state* S, double* v) {
    for (int k = -2; k <= S->k1 + 2; k++)
        for (int j = -2; j <= S->j1 + 2; j++) {
            for (int m = 0; m < 3; m++) v[LS(S->i1 + m, j, k)] = v[LS(1 + m, j, k)];
            for (int m = 0; m < 3; m++) v[LS(-2 + m, j, k)] = v[LS(S->imax - 2 + m, j, k)];
        }
    for (int k = -2; k <= S->k1 + 2; k++)
        for (int i = -2; i <= S->i1 + 2; i++) {
            for (int m = 0; m < 3; m++) v[LS(i, S->j1 + m, k)] = v[LS(i, 1 + m, k)];
            for (int m = 0; m < 3; m++) v[LS(i, -2 + m, k)] = v[LS(i, S->jmax - 2 + m, k)];
        }
}

/* boundsls, 'line' extrapolation of one ghost layer: bound.f90:386-483 */
static void boundsls(oracle_state* S, double* y) {
    for (int j = -2; j <= S->j1 + 2; j++)
        for (int i = -2; i <= S->i1 + 2; i++) {
            y[LS(i, j, 0)] = 2. * y[LS(i, j, 1)] - y[LS(i, j, 2)];
            y[LS(i, j, S->k1)] = 2. * y[LS(i, j, S->kmax)] - y[LS(i, j, S->kmax - 1)];
        }
    updthalos_ls(S, y);
}

/* boundmls, three extrapolated layers (init only): bound.f90:518-631 */
static void boundmls(oracle_state* S) {
    for (int l = 0; l < S->lmax; l++) {
        double* y = S->lvset[l];
        for (int j = -2; j <= S->j1 + 2; j++)
            for (int i = -2; i <= S->i1 + 2; i++) {
                y[LS(i, j, 0)] = 2. * y[LS(i, j, 1)] - y[LS(i, j, 2)];
                y[LS(i, j, -1)] = 2. * y[LS(i, j, 0)] - y[LS(i, j, 1)];
                y[LS(i, j, -2)] = 2. * y[LS(i, j, -1)] - y[LS(i, j, 0)];
                y[LS(i, j, S->k1)] = 2. * y[LS(i, j, S->kmax)] - y[LS(i, j, S->kmax - 1)];
                y[LS(i, j, S->k1 + 1)] = 2. * y[LS(i, j, S->k1)] - y[LS(i, j, S->kmax)];
                y[LS(i, j, S->k1 + 2)] = 2. * y[LS(i, j, S->k1 + 1)] - y[LS(i, j, S->k1)];
            }
    }
    for (int l = 0; l < S->lmax; l++) updthalos_ls(S, S->lvset[l]);
}

/* boundp: Neumann in z, periodic x/y (bound.f90:182-204) */
static void boundp_arr(oracle_state* S, double* p) {
    for (int j = 0; j <= S->j1; j++)
        for (int i = 0; i <= S->i1; i++) {
            p[CC(i, j, 0)] = p[CC(i, j, 1)];
            p[CC(i, j, S->k1)] = p[CC(i, j, S->kmax)];
        }
    updthalos(S, p);
}

/* boundc: periodic x/y only, z ghosts untouched (bound.f90:292-305) */
static void boundc(oracle_state* S, double* c) { updthalos(S, c); }

/* bounduvw, BC_in_z = 'Dirichl', BC_in_y = 'Periodic' (bound.f90:15-78) */
static void bounduvw_arr(oracle_state* S, double* u, double* v, double* w) {
    double v_lower, v_upper;
    if (S->P.ns_np) { v_lower = 0.0; v_upper = 0.0; }
    else { v_lower = -S->P.wall_speed; v_upper = S->P.wall_speed; }
    for (int j = 0; j <= S->j1; j++)
        for (int i = 0; i <= S->i1; i++) {
            u[CC(i, j, 0)] = -u[CC(i, j, 1)];
            u[CC(i, j, S->k1)] = -u[CC(i, j, S->kmax)];
            v[CC(i, j, 0)] = 2. * v_lower - v[CC(i, j, 1)];
            v[CC(i, j, S->k1)] = 2. * v_upper - v[CC(i, j, S->kmax)];
            w[CC(i, j, 0)] = 0.;
            w[CC(i, j, S->kmax)] = 0.;
            w[CC(i, j, S->k1)] = w[CC(i, j, S->kmax - 1)];
        }
    updthalos(S, u); updthalos(S, v); updthalos(S, w);
}

/* ---------------------------------------------------------------- FFT (VFFTPACK semantics) */
static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* unscaled complex DFT, sign = -1 forward, +1 backward */
static void cdft(int n, double* re, double* im, int sign) {
    const double argh = 2. * PI_F77 / (double)n;
    if (is_pow2(n)) {
        /* bit reversal */
        for (int i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) { double t = re[i]; re[i] = re[j]; re[j] = t; t = im[i]; im[i] = im[j]; im[j] = t; }
        }
        for (int len = 2; len <= n; len <<= 1) {
            int half = len >> 1, stride = n / len;
            for (int s = 0; s < n; s += len)
                for (int q = 0; q < half; q++) {
                    double ang = (double)(q * stride) * argh;
                    double wr = cos(ang), wi = (double)sign * sin(ang);
                    double xr = re[s + q + half], xi = im[s + q + half];
                    double tr = wr * xr - wi * xi, ti = wr * xi + wi * xr;
                    re[s + q + half] = re[s + q] - tr; im[s + q + half] = im[s + q] - ti;
                    re[s + q] = re[s + q] + tr; im[s + q] = im[s + q] + ti;
                }
        }
    } else {
        double* tr = (double*)malloc(sizeof(double) * n);
        double* ti = (double*)malloc(sizeof(double) * n);
        for (int k = 0; k < n; k++) {
            long double sr = 0, si = 0;
            for (int j = 0; j < n; j++) {
                double ang = (double)(((long)j * k) % n) * argh;
                double wr = cos(ang), wi = (double)sign * sin(ang);
                sr += (long double)wr * re[j] - (long double)wi * im[j];
                si += (long double)wr * im[j] + (long double)wi * re[j];
            }
            tr[k] = (double)sr; ti[k] = (double)si;
        }
        memcpy(re, tr, sizeof(double) * n); memcpy(im, ti, sizeof(double) * n);
        free(tr); free(ti);
    }
}

/* VRFFTF with M = 1 (fft.f:1073-1145, scale at 1770) */
void oracle_rfft_forward(int n, double* r) {
    double* re = (double*)malloc(sizeof(double) * n);
    double* im = (double*)calloc((size_t)n, sizeof(double));
    memcpy(re, r, sizeof(double) * n);
    cdft(n, re, im, -1);
    const double scale = sqrt(1. / (double)n);
    r[0] = scale * re[0];
    for (int k = 1; 2 * k < n; k++) { r[2 * k - 1] = scale * re[k]; r[2 * k] = scale * im[k]; }
    if (n % 2 == 0) r[n - 1] = scale * re[n / 2];
    free(re); free(im);
}

/* VRFFTB with M = 1 (fft.f:374-450, scale at 1057) */
void oracle_rfft_backward(int n, double* r) {
    double* re = (double*)calloc((size_t)n, sizeof(double));
    double* im = (double*)calloc((size_t)n, sizeof(double));
    re[0] = r[0];
    for (int k = 1; 2 * k < n; k++) {
        re[k] = r[2 * k - 1]; im[k] = r[2 * k];
        re[n - k] = r[2 * k - 1]; im[n - k] = -r[2 * k];
    }
    if (n % 2 == 0) re[n / 2] = r[n - 1];
    cdft(n, re, im, +1);
    const double scale = sqrt(1. / (double)n);
    for (int j = 0; j < n; j++) r[j] = scale * re[j];
    free(re); free(im);
}

/* ---------------------------------------------------------------- least squares (init.f90:1587-1650) */
void oracle_pseudo_inverse(double dx, double dy, double dz, double* out) {
    long double A[27][10], M[10][20];
    int cnt = 0;
    for (int dzn = -1; dzn <= 1; dzn++)
        for (int dyn = -1; dyn <= 1; dyn++)
            for (int dxn = -1; dxn <= 1; dxn++) {
                double a2 = dx * dxn, a3 = dy * dyn, a4 = dz * dzn;
                A[cnt][0] = 1.0L; A[cnt][1] = a2; A[cnt][2] = a3; A[cnt][3] = a4;
                A[cnt][4] = 0.5L * (long double)a2 * a2; A[cnt][5] = 0.5L * (long double)a3 * a3;
                A[cnt][6] = 0.5L * (long double)a4 * a4;
                A[cnt][7] = (long double)a2 * a3; A[cnt][8] = (long double)a2 * a4; A[cnt][9] = (long double)a3 * a4;
                cnt++;
            }
    for (int r = 0; r < 10; r++)
        for (int c = 0; c < 10; c++) {
            long double s = 0;
            for (int q = 0; q < 27; q++) s += A[q][r] * A[q][c];
            M[r][c] = s;
            M[r][10 + c] = (r == c) ? 1.0L : 0.0L;
        }
    for (int col = 0; col < 10; col++) {
        int piv = col;
        for (int r = col + 1; r < 10; r++) if (fabsl(M[r][col]) > fabsl(M[piv][col])) piv = r;
        if (piv != col) for (int c = 0; c < 20; c++) { long double t = M[col][c]; M[col][c] = M[piv][c]; M[piv][c] = t; }
        long double d = M[col][col];
        for (int c = 0; c < 20; c++) M[col][c] /= d;
        for (int r = 0; r < 10; r++) if (r != col) {
            long double f = M[r][col];
            if (f != 0) for (int c = 0; c < 20; c++) M[r][c] -= f * M[col][c];
        }
    }
    for (int r = 0; r < 10; r++)
        for (int q = 0; q < 27; q++) {
            long double s = 0;
            for (int c = 0; c < 10; c++) s += M[r][10 + c] * A[q][c];
            out[r + 10 * q] = (double)s;
        }
}

/* x = matmul(pseudo_inv_A, reshape(f(i-1:i+1,j-1:j+1,k-1:k+1),(/27/))), row r only
 * (interface.f90:59, 207); accumulation over the 27 columns in order. */
static inline double lsq_row(const double* A, int r, const double* v27) {
    double s = 0.;
    for (int q = 0; q < 27; q++) s = s + A[r + 10 * q] * v27[q];
    return s;
}

/* ---------------------------------------------------------------- pointwise helpers */
/* Heaviside 'trig' (misc.f90:439-459) */
double oracle_heaviside_trig(double phi, double eps) {
    const double pi = acos(-1.);
    if (phi < -eps) return 0.;
    else if (phi < eps) return .5 + .5 * phi / eps + .5 / pi * sin(pi * phi / eps);
    return 1.;
}
/* Dirac 'trig' (misc.f90:518-538) */
double oracle_dirac_trig(double phi, double eps) {
    const double pi = acos(-1.);
    if (phi < -eps) return 0.;
    else if (phi < eps) return .5 / eps + .5 / eps * cos(pi * phi / eps);
    return 0.;
}
/* icurvature, 3D branch (interface.f90:126-179); ph, cv indexed -1..1 via +1 */
double oracle_icurvature(int nb, const double* ph, const double* cv) {
    double sp = 0., icurv;
    if (cv[1] < 0.) sp = -1.;
    if (cv[1] > 0.) sp = 1.;
    double eps_0 = sp * ph[1];
    double eps_nb = sp * ph[1 + nb];
    if (cv[1] * cv[1 + nb] > 0.) {
        double a1 = eps_nb * cv[1];
        double a2 = eps_0 * cv[1 + nb];
        icurv = (a1 - a2) / (eps_nb - eps_0);
    } else {
        icurv = 0.;
    }
    return icurv;
}

/* Thomas elimination of one column (solver.f90:119-160) with the Periodic-branch
 * coefficients of initsolver.f90:36-49 */
void oracle_thomas(int kmax, double dzi, double lambda, double* pz) {
    double* a = (double*)malloc(sizeof(double) * kmax);
    double* b = (double*)malloc(sizeof(double) * kmax);
    double* c = (double*)malloc(sizeof(double) * kmax);
    double* d = (double*)malloc(sizeof(double) * kmax);
    for (int k = 0; k < kmax; k++) { a[k] = dzi * dzi; c[k] = dzi * dzi; b[k] = -(a[k] + c[k]); }
    b[0] = b[0] + a[0]; b[kmax - 1] = b[kmax - 1] + c[kmax - 1]; a[0] = 0.; c[kmax - 1] = 0.;
    double z = 1. / (b[0] + lambda);
    d[0] = c[0] * z;
    pz[0] = pz[0] * z;
    for (int k = 1; k < kmax - 1; k++) {
        double bb = b[k] + lambda;
        z = 1. / (bb - a[k] * d[k - 1]);
        d[k] = c[k] * z;
        pz[k] = (pz[k] - a[k] * pz[k - 1]) * z;
    }
    {
        double bb = b[kmax - 1] + lambda;
        z = bb - a[kmax - 1] * d[kmax - 2];
        if (z != 0.) pz[kmax - 1] = (pz[kmax - 1] - a[kmax - 1] * pz[kmax - 2]) / z;
        else pz[kmax - 1] = 0.;
    }
    for (int k = kmax - 2; k >= 0; k--) pz[k] = pz[k] - d[k] * pz[k + 1];
    free(a); free(b); free(c); free(d);
}

/* ---------------------------------------------------------------- create / destroy */
oracle_state* oracle_create(const icls_params* p) {
    if (!p || p->itot < 8 || p->jtot < 8 || p->ktot < 8 || p->lmax < 1 || p->lmax > ICLS_MAX_DROPS) return NULL;
    oracle_state* S = (oracle_state*)calloc(1, sizeof(oracle_state));
    S->P = *p;
    S->imax = p->itot; S->jmax = p->jtot; S->kmax = p->ktot;
    S->i1 = S->imax + 1; S->j1 = S->jmax + 1; S->k1 = S->kmax + 1;
    S->lmax = p->lmax;
    S->lsx = S->imax + 6; S->lsy = S->jmax + 6; S->lsz = S->kmax + 6;
    S->cx = S->imax + 2; S->cy = S->jmax + 2; S->cz = S->kmax + 2;
    S->nls = (long)S->lsx * S->lsy * S->lsz;
    S->nc = (long)S->cx * S->cy * S->cz;
    S->ni = (long)S->imax * S->jmax * S->kmax;
    /* param.f90:33-41, 63, 73 */
    S->lx = p->lx; S->ly = p->ly; S->lz = p->lz;
    S->dxi = p->itot / S->lx; S->dyi = p->jtot / S->ly; S->dzi = p->ktot / S->lz;
    S->dx = 1. / S->dxi; S->dy = 1. / S->dyi; S->dz = 1. / S->dzi;
    S->NB2 = 6. * S->dz;
    S->rho0 = p->rho1 < p->rho2 ? p->rho1 : p->rho2;
    S->visc = 1. / p->Re;
    S->pi = acos(-1.);
    S->dt = p->timestep;
#define AC(f) S->f = zalloc(S->nc)
    AC(unew); AC(vnew); AC(wnew); AC(pnew); AC(dudt); AC(dvdt); AC(dwdt);
    AC(jiemian); AC(miu); AC(rho); AC(u_l); AC(v_l); AC(w_l); AC(slset); AC(curv_cmn);
    AC(p_x); AC(p_y); AC(p_z); AC(p_xold); AC(p_yold); AC(p_zold); AC(pold); AC(p_jump); AC(p);
#undef AC
#define AI(f) S->f = zalloc(S->ni)
    AI(uccnew); AI(vccnew); AI(wccnew); AI(rho_u); AI(rho_v); AI(rho_w);
    AI(nrm_m); AI(nrm_x); AI(nrm_y); AI(nrm_z); AI(duold); AI(dvold); AI(dwold);
#undef AI
    S->hnew_mls = (double**)calloc((size_t)S->lmax, sizeof(double*));
    S->lvset = (double**)calloc((size_t)S->lmax, sizeof(double*));
    for (int l = 0; l < S->lmax; l++) { S->hnew_mls[l] = zalloc(S->nc); S->lvset[l] = zalloc(S->nls); }
    S->volume = zalloc(S->lmax); S->init_volume = zalloc(S->lmax);
    S->dmin = zalloc((long)S->lmax * S->lmax);
    S->a = zalloc(S->kmax); S->b = zalloc(S->kmax); S->c = zalloc(S->kmax);
    S->xyrt = zalloc((long)S->imax * S->jmax);
    for (long q = 0; q < S->ni; q++) S->nrm_m[q] = 1.;
    return S;
}

void oracle_destroy(oracle_state* S) {
    if (!S) return;
    double* arrs[] = {S->unew, S->vnew, S->wnew, S->pnew, S->dudt, S->dvdt, S->dwdt, S->jiemian, S->miu,
                      S->rho, S->u_l, S->v_l, S->w_l, S->slset, S->curv_cmn, S->p_x, S->p_y, S->p_z,
                      S->p_xold, S->p_yold, S->p_zold, S->pold, S->p_jump, S->p, S->uccnew, S->vccnew,
                      S->wccnew, S->rho_u, S->rho_v, S->rho_w, S->nrm_m, S->nrm_x, S->nrm_y, S->nrm_z,
                      S->duold, S->dvold, S->dwold, S->volume, S->init_volume, S->dmin, S->a, S->b, S->c, S->xyrt};
    for (size_t q = 0; q < sizeof(arrs) / sizeof(arrs[0]); q++) free(arrs[q]);
    for (int l = 0; l < S->lmax; l++) { free(S->hnew_mls[l]); free(S->lvset[l]); }
    free(S->hnew_mls); free(S->lvset);
    free(S->nb_i); free(S->nb_j); free(S->nb_k);
    free(S);
}

/* ---------------------------------------------------------------- initial conditions */
/* init_vel_p 'zer' / 'cou' (init.f90:148-158, 191-210) */
static void init_vel_p(oracle_state* S) {
    for (long q = 0; q < S->nc; q++) {
        S->unew[q] = 0.; S->vnew[q] = 0.; S->wnew[q] = 0.;
        S->dudt[q] = 0.; S->dvdt[q] = 0.; S->dwdt[q] = 0.; S->pnew[q] = 0.;
    }
    if (S->P.iniu == ICLS_INIU_COU) {
        double v_upper = S->P.wall_speed, v_lower = -S->P.wall_speed;
        double shear_rate = (v_upper - v_lower) / S->lz;
        for (int k = 0; k <= S->k1; k++) {
            double coorz = (k - 0.5) * S->dz;
            for (int j = 0; j <= S->j1; j++)
                for (int i = 0; i <= S->i1; i++) S->vnew[CC(i, j, k)] = v_lower + shear_rate * coorz;
        }
    }
}

/* init_levelset 'drp' / 'mdp': one sphere SDF per level set with the +lx and +ly
 * periodic images (init.f90:717-753, 945-980) */
static void init_levelset(oracle_state* S) {
    for (int l = 0; l < S->lmax; l++) {
        double xc = S->P.drop_xc[l], yc = S->P.drop_yc[l], zc = S->P.drop_zc[l], R = S->P.drop_r[l];
        double* ph = S->lvset[l];
        for (int k = -2; k <= S->k1 + 2; k++) {
            double coorz = (k - 0.5) / S->dzi;
            for (int j = -2; j <= S->j1 + 2; j++) {
                double coory = 0. + (j - 0.5) / S->dyi;
                for (int i = -2; i <= S->i1 + 2; i++) {
                    double coorx = 0. + (i - 0.5) / S->dxi;
                    double v = sqrt((coorx - xc) * (coorx - xc) + (coory - yc) * (coory - yc) + (coorz - zc) * (coorz - zc)) - R;
                    double p1 = sqrt((coorx - xc) * (coorx - xc) + (coory - yc - S->ly) * (coory - yc - S->ly) + (coorz - zc) * (coorz - zc)) - R;
                    if (v > 0.) v = fmin(v, p1);
                    double p2 = sqrt((coorx - xc - S->lx) * (coorx - xc - S->lx) + (coory - yc) * (coory - yc) + (coorz - zc) * (coorz - zc)) - R;
                    if (v > 0.) v = fmin(v, p2);
                    ph[LS(i, j, k)] = v;
                }
            }
        }
    }
}

/* initsolver, Periodic branch (initsolver.f90:26-83) */
static void initsolver(oracle_state* S) {
    for (int k = 0; k < S->kmax; k++) { S->a[k] = S->dzi * S->dzi; S->c[k] = S->dzi * S->dzi; S->b[k] = -(S->a[k] + S->c[k]); }
    S->b[0] = S->b[0] + S->a[0];
    S->b[S->kmax - 1] = S->b[S->kmax - 1] + S->c[S->kmax - 1];
    S->a[0] = 0.; S->c[S->kmax - 1] = 0.;
    int itot = S->imax, jtot = S->jmax;
    double* xrt = zalloc(itot + 1);
    double* yrt = zalloc(jtot + 1);
    for (int i = 3; i <= itot; i += 2) {
        double s = sin((double)(i - 1) * S->pi / (2. * itot));
        xrt[i - 1] = -4. * S->dxi * S->dxi * (s * s);
        xrt[i] = xrt[i - 1];
    }
    xrt[1] = 0.; xrt[itot] = -4. * S->dxi * S->dxi;
    for (int j = 3; j <= jtot; j += 2) {
        double s = sin((double)(j - 1) * S->pi / (2. * jtot));
        yrt[j - 1] = -4. * S->dyi * S->dyi * (s * s);
        yrt[j] = yrt[j - 1];
    }
    yrt[1] = 0.; yrt[jtot] = -4. * S->dyi * S->dyi;
    for (int j = 1; j <= S->jmax; j++)
        for (int i = 1; i <= S->imax; i++) S->xyrt[(i - 1) + (long)S->imax * (j - 1)] = xrt[i] + yrt[j];
    free(xrt); free(yrt);
}

/* ---------------------------------------------------------------- level set: tube + advection */
/* tube_construction (levelset.f90:87-197) */
int oracle_tube_construction(oracle_state* S, int l) {
    double* phi = S->lvset[l];
    const double NB2 = S->NB2;
    OMP_FOR
    for (long q = 0; q < S->nls; q++) {
        if (phi[q] > NB2) phi[q] = NB2;
        else if (phi[q] < -NB2) phi[q] = -NB2;
    }
    int cnt = 0;
    int* mask = (int*)malloc(sizeof(int) * S->ni);
    OMP_FOR_RED(+ : cnt)
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) {
                int m;
                if (fabs(phi[LS(i, j, k)]) < NB2) m = 1;
                else {
                    double xsum = 0., ysum = 0., zsum = 0.;
                    for (int q = -2; q <= 2; q++) xsum = xsum + (fabs(phi[LS(i + q, j, k)]) - NB2);
                    for (int q = -2; q <= 2; q++) ysum = ysum + (fabs(phi[LS(i, j + q, k)]) - NB2);
                    for (int q = -2; q <= 2; q++) zsum = zsum + (fabs(phi[LS(i, j, k + q)]) - NB2);
                    m = (xsum == 0. && ysum == 0. && zsum == 0.) ? 0 : 1;
                }
                mask[II(i, j, k)] = m;
                cnt += m;
            }
    S->nb_cnt = cnt;
    free(S->nb_i); free(S->nb_j); free(S->nb_k);
    S->nb_i = (int*)malloc(sizeof(int) * (cnt + 1));
    S->nb_j = (int*)malloc(sizeof(int) * (cnt + 1));
    S->nb_k = (int*)malloc(sizeof(int) * (cnt + 1));
    cnt = 0;
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++)
                if (mask[II(i, j, k)] == 1) { S->nb_i[cnt] = i; S->nb_j[cnt] = j; S->nb_k[cnt] = k; cnt++; }
    free(mask);
    boundsls(S, phi);
    return 0;
}

int oracle_band_count(oracle_state* S) { return S->nb_cnt; }
int oracle_band_mask(oracle_state* S, int* mask) {
    memset(mask, 0, sizeof(int) * S->ni);
    for (int c = 0; c < S->nb_cnt; c++) mask[II(S->nb_i[c], S->nb_j[c], S->nb_k[c])] = 1;
    return 0;
}

/* HOUC5_nb (hyperPDE.f90:397-495) */
static void houc5_nb(oracle_state* S, const double* vu, const double* vv, const double* vw, const double* phi, double* dfdt) {
    const double dx = S->dx, dy = S->dy, dz = S->dz;
    const int kmax = S->kmax;
    memset(dfdt, 0, sizeof(double) * S->ni);
    OMP_FOR
    for (int cnt = 0; cnt < S->nb_cnt; cnt++) {
        int i = S->nb_i[cnt], j = S->nb_j[cnt], k = S->nb_k[cnt];
        long c = II(i, j, k);
        double dfx, dfy, dfz, udfdx, vdfdy, wdfdz;
        if (vu[c] >= 0.)
            dfx = 1. / 60. * (-2. * phi[LS(i - 3, j, k)] + 15. * phi[LS(i - 2, j, k)] - 60. * phi[LS(i - 1, j, k)] +
                              20. * phi[LS(i, j, k)] + 30. * phi[LS(i + 1, j, k)] - 3. * phi[LS(i + 2, j, k)]);
        else
            dfx = 1. / 60. * (2. * phi[LS(i + 3, j, k)] - 15. * phi[LS(i + 2, j, k)] + 60. * phi[LS(i + 1, j, k)] -
                              20. * phi[LS(i, j, k)] - 30. * phi[LS(i - 1, j, k)] + 3. * phi[LS(i - 2, j, k)]);
        udfdx = -vu[c] * dfx / dx;
        if (vv[c] >= 0.)
            dfy = 1. / 60. * (-2. * phi[LS(i, j - 3, k)] + 15. * phi[LS(i, j - 2, k)] - 60. * phi[LS(i, j - 1, k)] +
                              20. * phi[LS(i, j, k)] + 30. * phi[LS(i, j + 1, k)] - 3. * phi[LS(i, j + 2, k)]);
        else
            dfy = 1. / 60. * (2. * phi[LS(i, j + 3, k)] - 15. * phi[LS(i, j + 2, k)] + 60. * phi[LS(i, j + 1, k)] -
                              20. * phi[LS(i, j, k)] - 30. * phi[LS(i, j - 1, k)] + 3. * phi[LS(i, j - 2, k)]);
        vdfdy = -vv[c] * dfy / dy;
        if (k > 2 && k < kmax - 1) {
            if (vw[c] >= 0.)
                dfz = 1. / 60. * (-2. * phi[LS(i, j, k - 3)] + 15. * phi[LS(i, j, k - 2)] - 60. * phi[LS(i, j, k - 1)] +
                                  20. * phi[LS(i, j, k)] + 30. * phi[LS(i, j, k + 1)] - 3. * phi[LS(i, j, k + 2)]);
            else
                dfz = 1. / 60. * (2. * phi[LS(i, j, k + 3)] - 15. * phi[LS(i, j, k + 2)] + 60. * phi[LS(i, j, k + 1)] -
                                  20. * phi[LS(i, j, k)] - 30. * phi[LS(i, j, k - 1)] + 3. * phi[LS(i, j, k - 2)]);
            wdfdz = -vw[c] * dfz / dz;
        } else if (k == 2 || k == kmax - 1) {
            if (vw[c] >= 0.)
                dfz = (1. / 6.) * phi[LS(i, j, k - 2)] - phi[LS(i, j, k - 1)] + (1. / 2.) * phi[LS(i, j, k)] + (1. / 3.) * phi[LS(i, j, k + 1)];
            else
                dfz = -(1. / 6.) * phi[LS(i, j, k + 2)] + phi[LS(i, j, k + 1)] - (1. / 2.) * phi[LS(i, j, k)] - (1. / 3.) * phi[LS(i, j, k - 1)];
            wdfdz = -vw[c] * dfz / dz;
        } else {
            if (vw[c] >= 0.) dfz = phi[LS(i, j, k)] - phi[LS(i, j, k - 1)];
            else dfz = phi[LS(i, j, k + 1)] - phi[LS(i, j, k)];
            wdfdz = -vw[c] * dfz / dz;
        }
        dfdt[c] = udfdx + vdfdy + wdfdz;
    }
}

/* one WENO5 upwind reconstruction from the five differences (hyperPDE.f90:551-576) */
static inline double weno5(double v1, double v2, double v3, double v4, double v5) {
    const double eps_weno = 1.e-6, sigma_1 = 0.1, sigma_2 = 0.6, sigma_3 = 0.3;
    double S1 = (13. / 12.) * ((v1 - 2. * v2 + v3) * (v1 - 2. * v2 + v3)) + (1. / 4.) * ((v1 - 4. * v2 + 3. * v3) * (v1 - 4. * v2 + 3. * v3));
    double S2 = (13. / 12.) * ((v2 - 2. * v3 + v4) * (v2 - 2. * v3 + v4)) + (1. / 4.) * ((v2 - v4) * (v2 - v4));
    double S3 = (13. / 12.) * ((v3 - 2. * v4 + v5) * (v3 - 2. * v4 + v5)) + (1. / 4.) * ((3. * v3 - 4. * v4 + v5) * (3. * v3 - 4. * v4 + v5));
    double a1 = sigma_1 / ((eps_weno + S1) * (eps_weno + S1));
    double a2 = sigma_2 / ((eps_weno + S2) * (eps_weno + S2));
    double a3 = sigma_3 / ((eps_weno + S3) * (eps_weno + S3));
    double w1 = a1 / (a1 + a2 + a3);
    double w2 = a2 / (a1 + a2 + a3);
    double w3 = a3 / (a1 + a2 + a3);
    return w1 * ((1. / 3.) * v1 - (7. / 6.) * v2 + (11. / 6.) * v3) +
           w2 * ((-1. / 6.) * v2 + (5. / 6.) * v3 + (1. / 3.) * v4) +
           w3 * ((1. / 3.) * v3 + (5. / 6.) * v4 - (1. / 6.) * v5);
}

/* left-biased (minus) and right-biased (plus) WENO5 derivatives along a line through
 * f[-3..3] with spacing h (hyperPDE.f90:545-549 and 579-583) */
static inline double weno_minus(const double* f, double h) {
    return weno5((f[-2] - f[-3]) / h, (f[-1] - f[-2]) / h, (f[0] - f[-1]) / h, (f[1] - f[0]) / h, (f[2] - f[1]) / h);
}
static inline double weno_plus(const double* f, double h) {
    return weno5((f[3] - f[2]) / h, (f[2] - f[1]) / h, (f[1] - f[0]) / h, (f[0] - f[-1]) / h, (f[-1] - f[-2]) / h);
}

/* WENO5_nb (hyperPDE.f90:503-800); y differences divided by dx as in the reference */
static void weno5_nb(oracle_state* S, const double* vu, const double* vv, const double* vw, const double* phi, double* dfdt) {
    const double dx = S->dx, dz = S->dz;
    const int kmax = S->kmax;
    memset(dfdt, 0, sizeof(double) * S->ni);
    OMP_FOR
    for (int cnt = 0; cnt < S->nb_cnt; cnt++) {
        int i = S->nb_i[cnt], j = S->nb_j[cnt], k = S->nb_k[cnt];
        long c = II(i, j, k);
        double f[7], dfx, dfy, dfz, udfdx, vdfdy, wdfdz;
        for (int q = -3; q <= 3; q++) f[q + 3] = phi[LS(i + q, j, k)];
        dfx = (vu[c] >= 0.) ? weno_minus(f + 3, dx) : weno_plus(f + 3, dx);
        udfdx = -vu[c] * dfx;
        for (int q = -3; q <= 3; q++) f[q + 3] = phi[LS(i, j + q, k)];
        dfy = (vv[c] >= 0.) ? weno_minus(f + 3, dx) : weno_plus(f + 3, dx);
        vdfdy = -vv[c] * dfy;
        if (k > 2 && k < kmax - 1) {
            for (int q = -3; q <= 3; q++) f[q + 3] = phi[LS(i, j, k + q)];
            dfz = (vw[c] >= 0.) ? weno_minus(f + 3, dz) : weno_plus(f + 3, dz);
            wdfdz = -vw[c] * dfz;
        } else if (k == 2 || k == kmax - 1) {
            if (vw[c] >= 0.)
                dfz = (1. / 6.) * phi[LS(i, j, k - 2)] - phi[LS(i, j, k - 1)] + (1. / 2.) * phi[LS(i, j, k)] + (1. / 3.) * phi[LS(i, j, k + 1)];
            else
                dfz = -(1. / 6.) * phi[LS(i, j, k + 2)] + phi[LS(i, j, k + 1)] - (1. / 2.) * phi[LS(i, j, k)] - (1. / 3.) * phi[LS(i, j, k - 1)];
            wdfdz = -vw[c] * dfz / dz;
        } else {
            if (vw[c] >= 0.) dfz = phi[LS(i, j, k)] - phi[LS(i, j, k - 1)];
            else dfz = phi[LS(i, j, k + 1)] - phi[LS(i, j, k)];
            wdfdz = -vw[c] * dfz / dz;
        }
        dfdt[c] = udfdx + vdfdy + wdfdz;
    }
}

static void adv_rhs(oracle_state* S, int scheme, const double* vu, const double* vv, const double* vw, const double* phi, double* dfdt) {
    if (scheme == ICLS_ADV_WENO5) weno5_nb(S, vu, vv, vw, phi, dfdt);
    else houc5_nb(S, vu, vv, vw, phi, dfdt);
}

/* RK_nb (levelset.f90:342-367) */
static void rk_nb(oracle_state* S, double* out, const double* in, double factor, const double* dfdt) {
    memcpy(out, in, sizeof(double) * S->nls);
    OMP_FOR
    for (int cnt = 0; cnt < S->nb_cnt; cnt++) {
        int i = S->nb_i[cnt], j = S->nb_j[cnt], k = S->nb_k[cnt];
        out[LS(i, j, k)] = in[LS(i, j, k)] + factor * S->dt * dfdt[II(i, j, k)];
    }
    boundsls(S, out);
}

/* SSP-RK3 over the band with a given advection scheme and velocity
 * (levelset.f90:235-243 / 268-276 / 555-565) */
static void rk3_advect(oracle_state* S, int scheme, const double* vu, const double* vv, const double* vw, double* phi_io) {
    long n = S->nls;
    double* phi_0 = phi_io;
    double* phi_1 = zalloc(n);
    double* phi_2 = zalloc(n);
    double* tmp = zalloc(n);
    double* dfdt = zalloc(S->ni);
    adv_rhs(S, scheme, vu, vv, vw, phi_0, dfdt);
    rk_nb(S, phi_1, phi_0, 1., dfdt);
    adv_rhs(S, scheme, vu, vv, vw, phi_1, dfdt);
    OMP_FOR
    for (long q = 0; q < n; q++) tmp[q] = 3. / 4. * phi_0[q] + 1. / 4. * phi_1[q];
    rk_nb(S, phi_2, tmp, 1. / 4., dfdt);
    adv_rhs(S, scheme, vu, vv, vw, phi_2, dfdt);
    OMP_FOR
    for (long q = 0; q < n; q++) tmp[q] = 1. / 3. * phi_0[q] + 2. / 3. * phi_2[q];
    rk_nb(S, phi_io, tmp, 2. / 3., dfdt);  /* phi_3 -> lvset */
    free(phi_1); free(phi_2); free(tmp); free(dfdt);
}

/* ls_advection (levelset.f90:208-335) */
int oracle_ls_advection(oracle_state* S, int l) {
    rk3_advect(S, S->P.ls_adv_space, S->uccnew, S->vccnew, S->wccnew, S->lvset[l]);
    return 0;
}

int oracle_eval_adv(oracle_state* S, int l, int scheme, double* dfdt) {
    adv_rhs(S, scheme, S->uccnew, S->vccnew, S->wccnew, S->lvset[l], dfdt);
    return 0;
}

/* ---------------------------------------------------------------- curvature (interface.f90) */
static void gather27(oracle_state* S, const double* f, int is_ls, int i, int j, int k, double* v) {
    int q = 0;
    for (int c = -1; c <= 1; c++)
        for (int b = -1; b <= 1; b++)
            for (int a = -1; a <= 1; a++) v[q++] = is_ls ? f[LS(i + a, j + b, k + c)] : f[CC(i + a, j + b, k + c)];
}

/* get_curvature (interface.f90:20-119) on an LS array into a C array */
static void get_curvature_arr(oracle_state* S, const double* phi, double* curvature) {
    const double dz = S->dz;
    OMP_FOR
    for (long q = 0; q < S->ni; q++) { S->nrm_m[q] = 1.; S->nrm_x[q] = 0.; S->nrm_y[q] = 0.; S->nrm_z[q] = 0.; }
    OMP_FOR
    for (long q = 0; q < S->nc; q++) curvature[q] = 0.;
    const double max_kappa = 1. / dz;
    OMP_FOR
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) {
                if (fabs(phi[LS(i, j, k)]) <= S->NB2 - 1. * dz) {
                    double v[27], x[10];
                    gather27(S, phi, 1, i, j, k, v);
                    for (int r = 0; r < 10; r++) x[r] = lsq_row(S->pinvA, r, v);
                    double dphidx = x[1], dphidy = x[2], dphidz = x[3];
                    double d2phidx = x[4], d2phidy = x[5], d2phidz = x[6];
                    double d2phidxdy = x[7], d2phidxdz = x[8], d2phidydz = x[9];
                    double norm_val = sqrt(dphidx * dphidx + dphidy * dphidy + dphidz * dphidz);
                    long c = II(i, j, k);
                    S->nrm_m[c] = norm_val;
                    S->nrm_x[c] = dphidx / norm_val;
                    S->nrm_y[c] = dphidy / norm_val;
                    S->nrm_z[c] = dphidz / norm_val;
                    double cv;
                    if (norm_val < 0.9) cv = 0.;
                    else {
                        double term1 = (d2phidx + d2phidy + d2phidz) / norm_val;
                        double n3 = norm_val * norm_val * norm_val;
                        double term2 = dphidx * dphidx * d2phidx + dphidy * dphidy * d2phidy + dphidz * dphidz * d2phidz;
                        term2 = term2 / n3;
                        double term3 = dphidx * dphidy * d2phidxdy + dphidx * dphidz * d2phidxdz + dphidy * dphidz * d2phidydz;
                        term3 = 2. * term3 / n3;
                        cv = -(term1 - term2 - term3);
                    }
                    if (cv < -max_kappa) cv = -max_kappa;
                    else if (cv > max_kappa) cv = max_kappa;
                    else if (fabs(cv) < 1e-9) cv = 0.;
                    curvature[CC(i, j, k)] = cv;
                }
            }
}

/* re_curvature (interface.f90:186-218) */
static void re_curvature_arr(oracle_state* S, const double* phi, const double* curv, double* re_curv) {
    OMP_FOR
    for (long q = 0; q < S->nc; q++) re_curv[q] = 0.;
    OMP_FOR
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++)
                if (fabs(phi[LS(i, j, k)]) < 2. * S->dz) {
                    double v[27];
                    gather27(S, curv, 0, i, j, k, v);
                    re_curv[CC(i, j, k)] = lsq_row(S->pinvA, 0, v);
                }
}

int oracle_get_curvature(oracle_state* S, int l) {
    get_curvature_arr(S, S->lvset[l], S->curv_cmn);
    boundc(S, S->curv_cmn);
    return 0;
}
int oracle_re_curvature(oracle_state* S, int l) {
    double* re = zalloc(S->nc);
    re_curvature_arr(S, S->lvset[l], S->curv_cmn, re);
    memcpy(S->curv_cmn, re, sizeof(double) * S->nc);
    boundc(S, S->curv_cmn);
    free(re);
    return 0;
}

/* ---------------------------------------------------------------- mass correction */
/* inflation, ispeed_dependence = 'curv', i_space = 'HOUC5' (levelset.f90:407-582) */
int oracle_inflation(oracle_state* S, int l) {
    const double s1 = 1.5 * S->dz;
    double* kappa = zalloc(S->nc);
    double* dheav = zalloc(S->nc);
    double* delta = zalloc(S->ni);
    double* u_lc = zalloc(S->ni);
    double* v_lc = zalloc(S->ni);
    double* w_lc = zalloc(S->ni);
    double* phi = S->lvset[l];
    get_curvature_arr(S, phi, kappa);
    boundc(S, kappa);
    OMP_FOR
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) delta[II(i, j, k)] = oracle_dirac_trig(phi[LS(i, j, k)], s1);
    OMP_FOR
    for (int k = 0; k <= S->k1; k++)
        for (int j = 0; j <= S->j1; j++)
            for (int i = 0; i <= S->i1; i++) dheav[CC(i, j, k)] = oracle_heaviside_trig(phi[LS(i, j, k)], s1);
    /* area: sum(delta*kappa*normal%m)*dx*dy*dz (levelset.f90:499) */
    double temp = 0.;
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) temp = temp + delta[II(i, j, k)] * kappa[CC(i, j, k)] * S->nrm_m[II(i, j, k)];
    double area = temp * S->dx * S->dy * S->dz;
    if (S->P.mass_correction == 2) {
        /* normalised variant (not in the reference): the H-measured volume changes at
         * temp * sum(delta^2 kappa |grad phi|^2) dV under the correction velocity */
        temp = 0.;
        for (int k = 1; k <= S->kmax; k++)
            for (int j = 1; j <= S->jmax; j++)
                for (int i = 1; i <= S->imax; i++) {
                    const double d = delta[II(i, j, k)], m = S->nrm_m[II(i, j, k)];
                    temp = temp + d * d * kappa[CC(i, j, k)] * (m * m);
                }
        area = temp * S->dx * S->dy * S->dz;
    }
    temp = 0.;
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) temp = temp + dheav[CC(i, j, k)];
    S->volume[l] = 1. - temp / (1. * S->imax * S->jmax * S->kmax);
    double loss_rate = S->lx * S->ly * S->lz * (S->volume[l] - S->init_volume[l]) / (S->P.inflate_step * S->dt);
    temp = loss_rate / area;
    S->maxchk_glb = temp / s1;
    boundp_arr(S, dheav);
    OMP_FOR
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) {
                S->u_l[CC(i, j, k)] = -temp * (dheav[CC(i + 1, j, k)] - dheav[CC(i, j, k)]) / S->dx;
                S->v_l[CC(i, j, k)] = -temp * (dheav[CC(i, j + 1, k)] - dheav[CC(i, j, k)]) / S->dy;
                S->w_l[CC(i, j, k)] = -temp * (dheav[CC(i, j, k + 1)] - dheav[CC(i, j, k)]) / S->dz;
            }
    bounduvw_arr(S, S->u_l, S->v_l, S->w_l);
    OMP_FOR
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) {
                long c = II(i, j, k);
                u_lc[c] = (S->u_l[CC(i - 1, j, k)] + S->u_l[CC(i, j, k)]) / 2. * kappa[CC(i, j, k)];
                v_lc[c] = (S->v_l[CC(i, j - 1, k)] + S->v_l[CC(i, j, k)]) / 2. * kappa[CC(i, j, k)];
                w_lc[c] = (S->w_l[CC(i, j, k - 1)] + S->w_l[CC(i, j, k)]) / 2. * kappa[CC(i, j, k)];
            }
    rk3_advect(S, ICLS_ADV_HOUC5, u_lc, v_lc, w_lc, phi);
    free(kappa); free(dheav); free(delta); free(u_lc); free(v_lc); free(w_lc);
    return 1;
}

/* ls_mass_correction (levelset.f90:377-400); returns 1 when inflation ran */
int oracle_ls_mass_correction(oracle_state* S, int istep, int l) {
    S->maxchk_glb = 0.;
    if (istep % S->P.inflate_step == 0) {
        double massloss = (S->volume[l] - S->init_volume[l]) / S->init_volume[l];
        /* levelset.f90:392; mass_trigger_exp > 0 lowers the 1e-5 (an extension) */
        const double trig = S->P.mass_trigger_exp > 0 ? pow(10., -S->P.mass_trigger_exp) : 1e-5;
        if (fabs(massloss) > trig) return oracle_inflation(S, l);
    }
    return 0;
}

/* ---------------------------------------------------------------- re-initialisation */
/* WENO5_deri + WENO5_reinit (hyperPDE.f90:2646-2746, 2932-3197) */
static void weno5_reinit(oracle_state* S, const double* sgn, const double* phi, double* RHS) {
    const double dx = S->dx, dz = S->dz, lim = S->NB2 + dz;
    const int kmax = S->kmax;
    OMP_FOR
    for (long q = 0; q < S->nls; q++) RHS[q] = 0.;
    OMP_FOR
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) {
                if (!(fabs(phi[LS(i, j, k)]) < lim)) continue;
                double f[7], A, B, C, D, E, F;
                for (int q = -3; q <= 3; q++) f[q + 3] = phi[LS(i + q, j, k)];
                A = weno_minus(f + 3, dx); B = weno_plus(f + 3, dx);
                for (int q = -3; q <= 3; q++) f[q + 3] = phi[LS(i, j + q, k)];
                C = weno_minus(f + 3, dx); D = weno_plus(f + 3, dx);
                if (k >= 3 && k <= kmax - 2) {
                    for (int q = -3; q <= 3; q++) f[q + 3] = phi[LS(i, j, k + q)];
                    E = weno_minus(f + 3, dz); F = weno_plus(f + 3, dz);
                } else {
                    E = (phi[LS(i, j, k)] - phi[LS(i, j, k - 1)]) / dz;
                    F = (phi[LS(i, j, k + 1)] - phi[LS(i, j, k)]) / dz;
                }
                double s = sgn[II(i, j, k)], dydx, dydy, dydz;
                if ((B * s < 0.) && ((A + B) * s < 0.)) dydx = B;
                else if ((A * s > 0.) && ((A + B) * s > 0.)) dydx = A;
                else dydx = 0.5 * (A + B);
                if ((D * s < 0.) && ((C + D) * s < 0.)) dydy = D;
                else if ((C * s > 0.) && ((C + D) * s > 0.)) dydy = C;
                else dydy = 0.5 * (C + D);
                if ((F * s < 0.) && ((E + F) * s < 0.)) dydz = F;
                else if ((E * s > 0.) && ((E + F) * s > 0.)) dydz = E;
                else dydz = 0.5 * (E + F);
                RHS[LS(i, j, k)] = s * (1. - sqrt(dydx * dydx + dydy * dydy + dydz * dydz));
            }
}

static void reinit_sgn(oracle_state* S, const double* phi_0, double* sgn) {
    const double dz = S->dz;
    OMP_FOR
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) {
                double p = phi_0[LS(i, j, k)], s;
                if (p < -dz) s = -1.;
                else if (p > dz) s = 1.;
                else s = p / sqrt(p * p + dz * dz);
                sgn[II(i, j, k)] = s;
            }
}

int oracle_eval_reinit_rhs(oracle_state* S, int l, double* rhs) {
    double* sgn = zalloc(S->ni);
    reinit_sgn(S, S->lvset[l], sgn);
    weno5_reinit(S, sgn, S->lvset[l], rhs);
    free(sgn);
    return 0;
}

/* ls_reinit, WENO5 + RK3 (levelset.f90:619-882) */
int oracle_ls_reinit(oracle_state* S, int l, int maxiter, int* iter_out, double* residual_out) {
    long n = S->nls;
    double* phi_0 = zalloc(n);
    double* phi_n1 = zalloc(n);
    double* phi_n2 = zalloc(n);
    double* phi_new = zalloc(n);
    double* RHS = zalloc(n);
    double* sgn = zalloc(S->ni);
    memcpy(phi_0, S->lvset[l], sizeof(double) * n);
    memcpy(phi_new, phi_0, sizeof(double) * n);
    reinit_sgn(S, phi_0, sgn);
    const double dtau = 0.5 * S->dz;
    int iter = maxiter;
    double residual = 0.;
    for (int it = 1; it <= maxiter; it++) {
        weno5_reinit(S, sgn, phi_0, RHS);
        OMP_FOR
        for (long q = 0; q < n; q++) phi_n1[q] = phi_0[q] + dtau * RHS[q];
        boundsls(S, phi_n1);
        weno5_reinit(S, sgn, phi_n1, RHS);
        OMP_FOR
        for (long q = 0; q < n; q++) phi_n2[q] = 3. / 4. * phi_0[q] + 1. / 4. * (phi_n1[q] + dtau * RHS[q]);
        boundsls(S, phi_n2);
        weno5_reinit(S, sgn, phi_n2, RHS);
        OMP_FOR
        for (long q = 0; q < n; q++) phi_new[q] = 1. / 3. * phi_0[q] + 2. / 3. * (phi_n2[q] + dtau * RHS[q]);
        boundsls(S, phi_new);
        residual = 0.;
        OMP_FOR_RED(max : residual)
        for (long q = 0; q < n; q++) { double d = fabs(phi_new[q] - phi_0[q]); if (d > residual) residual = d; }
        if (residual < pow(S->dz, 5)) { iter = it; break; }
        iter = maxiter;
        memcpy(phi_0, phi_new, sizeof(double) * n);
    }
    memcpy(S->lvset[l], phi_new, sizeof(double) * n);
    if (iter_out) *iter_out = iter;
    if (residual_out) *residual_out = residual;
    free(phi_0); free(phi_n1); free(phi_n2); free(phi_new); free(RHS); free(sgn);
    return 0;
}

/* ---------------------------------------------------------------- materials (misc.f90:79-290) */
int oracle_update_materials(oracle_state* S) {
    const double eps = 1.5 * S->dz;
    double* h_rho = zalloc(S->nc);
    OMP_FOR
    for (int k = 0; k <= S->k1; k++)
        for (int j = 0; j <= S->j1; j++)
            for (int i = 0; i <= S->i1; i++) {
                double s = S->lvset[0][LS(i, j, k)];
                for (int l = 1; l < S->lmax; l++) s = fmin(s, S->lvset[l][LS(i, j, k)]);
                S->slset[CC(i, j, k)] = s;
            }
    for (int l = 0; l < S->lmax; l++) {
        OMP_FOR
        for (int k = 0; k <= S->k1; k++)
            for (int j = 0; j <= S->j1; j++)
                for (int i = 0; i <= S->i1; i++)
                    S->hnew_mls[l][CC(i, j, k)] = oracle_heaviside_trig(S->lvset[l][LS(i, j, k)], eps);
    }
    OMP_FOR
    for (long q = 0; q < S->nc; q++) {
        double h = S->hnew_mls[0][q];
        for (int l = 1; l < S->lmax; l++) h = fmin(h, S->hnew_mls[l][q]);
        S->jiemian[q] = h;
    }
    OMP_FOR
    for (long q = 0; q < S->nc; q++) S->miu[q] = S->P.miu1 * S->jiemian[q] + S->P.miu2 * (1. - S->jiemian[q]);
    for (int l = 0; l < S->lmax; l++) {
        double s = 0.;
        const double* h = S->hnew_mls[l];
        for (int k = 1; k <= S->kmax; k++)
            for (int j = 1; j <= S->jmax; j++)
                for (int i = 1; i <= S->imax; i++) s = s + h[CC(i, j, k)];
        S->volume[l] = 1. - s / (1. * S->imax * S->jmax * S->kmax);
        OMP_FOR
        for (long q = 0; q < S->nc; q++) h_rho[q] = (l == 0) ? h[q] : fmin(h_rho[q], h[q]);
    }
    OMP_FOR
    for (long q = 0; q < S->nc; q++) S->rho[q] = S->P.rho1 * h_rho[q] + S->P.rho2 * (1. - h_rho[q]);
    double sv = 0.;
    for (int l = 0; l < S->lmax; l++) sv = sv + S->volume[l];
    S->rho_avr = S->P.rho1 * (1. - sv) + S->P.rho2 * sv;
    if (S->P.rho1 == S->P.rho2) {
        OMP_FOR
        for (long q = 0; q < S->ni; q++) { S->rho_u[q] = S->P.rho1; S->rho_v[q] = S->P.rho1; S->rho_w[q] = S->P.rho1; }
    } else {
        OMP_FOR
        for (int k = 1; k <= S->kmax; k++)
            for (int j = 1; j <= S->jmax; j++)
                for (int i = 1; i <= S->imax; i++) {
                    long c = II(i, j, k);
                    S->rho_u[c] = (S->rho[CC(i, j, k)] + S->rho[CC(i + 1, j, k)]) / 2.;
                    S->rho_v[c] = (S->rho[CC(i, j, k)] + S->rho[CC(i, j + 1, k)]) / 2.;
                    S->rho_w[c] = (S->rho[CC(i, j, k)] + S->rho[CC(i, j, k + 1)]) / 2.;
                }
    }
    free(h_rho);
    return 0;
}

/* ---------------------------------------------------------------- ls_main (levelset.f90:28-76) */
int oracle_ls_main(oracle_state* S, int istep) {
    for (int l = 0; l < S->lmax; l++) {
        oracle_tube_construction(S, l);
        oracle_ls_advection(S, l);
        if (S->P.mass_correction) oracle_ls_mass_correction(S, istep, l);
        if (S->P.re_initialization && istep % S->P.reinit_step == 0) {
            int iter; double res;
            oracle_ls_reinit(S, l, S->P.reinit_iter_max, &iter, &res);
            boundsls(S, S->lvset[l]);
            S->reinit_residual = res;
        }
    }
    oracle_update_materials(S);
    return 0;
}

/* ---------------------------------------------------------------- momentum (mom.f90) */
/* CNTR2_conv (mom.f90:67-130) */
static void cntr2_conv(oracle_state* S, double* RCu, double* RCv, double* RCw) {
    const double *u = S->unew, *v = S->vnew, *w = S->wnew;
    const double dx = S->dx, dy = S->dy, dz = S->dz;
    OMP_FOR
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) {
                int ip = i + 1, jp = j + 1, kp = k + 1, im = i - 1, jm = j - 1, km = k - 1;
                double uuip = 0.25 * (u[CC(ip, j, k)] + u[CC(i, j, k)]) * (u[CC(ip, j, k)] + u[CC(i, j, k)]);
                double uuim = 0.25 * (u[CC(im, j, k)] + u[CC(i, j, k)]) * (u[CC(im, j, k)] + u[CC(i, j, k)]);
                double uvjp = 0.25 * (u[CC(i, jp, k)] + u[CC(i, j, k)]) * (v[CC(ip, j, k)] + v[CC(i, j, k)]);
                double uvjm = 0.25 * (u[CC(i, jm, k)] + u[CC(i, j, k)]) * (v[CC(ip, jm, k)] + v[CC(i, jm, k)]);
                double uwkp = 0.25 * (u[CC(i, j, kp)] + u[CC(i, j, k)]) * (w[CC(ip, j, k)] + w[CC(i, j, k)]);
                double uwkm = 0.25 * (u[CC(i, j, km)] + u[CC(i, j, k)]) * (w[CC(ip, j, km)] + w[CC(i, j, km)]);
                long c = II(i, j, k);
                RCu[c] = (-uuip + uuim) / dx + (-uvjp + uvjm) / dy + (-uwkp + uwkm) / dz;
                double uvip = 0.25 * (u[CC(i, j, k)] + u[CC(i, jp, k)]) * (v[CC(i, j, k)] + v[CC(ip, j, k)]);
                double uvim = 0.25 * (u[CC(im, j, k)] + u[CC(im, jp, k)]) * (v[CC(i, j, k)] + v[CC(im, j, k)]);
                double vvjp = 0.25 * (v[CC(i, j, k)] + v[CC(i, jp, k)]) * (v[CC(i, j, k)] + v[CC(i, jp, k)]);
                double vvjm = 0.25 * (v[CC(i, j, k)] + v[CC(i, jm, k)]) * (v[CC(i, j, k)] + v[CC(i, jm, k)]);
                double wvkp = 0.25 * (w[CC(i, j, k)] + w[CC(i, jp, k)]) * (v[CC(i, j, kp)] + v[CC(i, j, k)]);
                double wvkm = 0.25 * (w[CC(i, j, km)] + w[CC(i, jp, km)]) * (v[CC(i, j, km)] + v[CC(i, j, k)]);
                RCv[c] = (-uvip + uvim) / dx + (-vvjp + vvjm) / dy + (-wvkp + wvkm) / dz;
                double uwip = 0.25 * (w[CC(i, j, k)] + w[CC(ip, j, k)]) * (u[CC(i, j, k)] + u[CC(i, j, kp)]);
                double uwim = 0.25 * (w[CC(i, j, k)] + w[CC(im, j, k)]) * (u[CC(im, j, k)] + u[CC(im, j, kp)]);
                double vwjp = 0.25 * (w[CC(i, j, k)] + w[CC(i, jp, k)]) * (v[CC(i, j, k)] + v[CC(i, j, kp)]);
                double vwjm = 0.25 * (w[CC(i, j, k)] + w[CC(i, jm, k)]) * (v[CC(i, jm, k)] + v[CC(i, jm, kp)]);
                double wwkp = 0.25 * (w[CC(i, j, k)] + w[CC(i, j, kp)]) * (w[CC(i, j, k)] + w[CC(i, j, kp)]);
                double wwkm = 0.25 * (w[CC(i, j, k)] + w[CC(i, j, km)]) * (w[CC(i, j, k)] + w[CC(i, j, km)]);
                RCw[c] = (-uwip + uwim) / dx + (-vwjp + vwjm) / dy + (-wwkp + wwkm) / dz;
            }
}

/* delta_diff (mom.f90:134-273) */
static void delta_diff(oracle_state* S, double* RDu, double* RDv, double* RDw) {
    const double *u = S->unew, *v = S->vnew, *w = S->wnew, *miu = S->miu;
    const double dx = S->dx, dy = S->dy, dz = S->dz;
    OMP_FOR
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) {
                int ip = i + 1, jp = j + 1, kp = k + 1, im = i - 1, jm = j - 1, km = k - 1;
                long c = II(i, j, k);
                double dudxp = (u[CC(ip, j, k)] - u[CC(i, j, k)]) / dx;
                double dudxm = (u[CC(i, j, k)] - u[CC(im, j, k)]) / dx;
                double dudyp = (u[CC(i, jp, k)] - u[CC(i, j, k)]) / dy;
                double dudym = (u[CC(i, j, k)] - u[CC(i, jm, k)]) / dy;
                double dvdxp = (v[CC(ip, j, k)] - v[CC(i, j, k)]) / dx;
                double dvdxm = (v[CC(ip, jm, k)] - v[CC(i, jm, k)]) / dx;
                double dudzp = (u[CC(i, j, kp)] - u[CC(i, j, k)]) / dz;
                double dudzm = (u[CC(i, j, k)] - u[CC(i, j, km)]) / dz;
                double dwdxp = (w[CC(ip, j, k)] - w[CC(i, j, k)]) / dx;
                double dwdxm = (w[CC(ip, j, km)] - w[CC(i, j, km)]) / dx;
                double ma1 = 0.25 * (miu[CC(i, j, k)] + miu[CC(i, jm, k)] + miu[CC(ip, jm, k)] + miu[CC(ip, j, k)]);
                double ma2 = 0.25 * (miu[CC(i, j, k)] + miu[CC(i, jp, k)] + miu[CC(ip, jp, k)] + miu[CC(ip, j, k)]);
                double ma3 = 0.25 * (miu[CC(i, j, k)] + miu[CC(i, j, km)] + miu[CC(ip, j, km)] + miu[CC(ip, j, k)]);
                double ma4 = 0.25 * (miu[CC(i, j, k)] + miu[CC(i, j, kp)] + miu[CC(ip, j, kp)] + miu[CC(ip, j, k)]);
                double RD1 = 2. * (dudxp * miu[CC(ip, j, k)] - dudxm * miu[CC(i, j, k)]) / dx;
                double RD2 = (dudyp * ma2 - dudym * ma1) / dy;
                double RD3 = (dvdxp * ma2 - dvdxm * ma1) / dy;
                double RD4 = (dudzp * ma4 - dudzm * ma3) / dz;
                double RD5 = (dwdxp * ma4 - dwdxm * ma3) / dz;
                RDu[c] = (RD1 + RD2 + RD3 + RD4 + RD5) / S->rho_u[c];

                dvdxp = (v[CC(ip, j, k)] - v[CC(i, j, k)]) / dx;
                dvdxm = (v[CC(i, j, k)] - v[CC(im, j, k)]) / dx;
                dudyp = (u[CC(i, jp, k)] - u[CC(i, j, k)]) / dy;
                dudym = (u[CC(im, jp, k)] - u[CC(im, j, k)]) / dy;
                double dvdyp = (v[CC(i, jp, k)] - v[CC(i, j, k)]) / dy;
                double dvdym = (v[CC(i, j, k)] - v[CC(i, jm, k)]) / dy;
                double dvdzp = (v[CC(i, j, kp)] - v[CC(i, j, k)]) / dz;
                double dvdzm = (v[CC(i, j, k)] - v[CC(i, j, km)]) / dz;
                double dwdyp = (w[CC(i, jp, k)] - w[CC(i, j, k)]) / dy;
                double dwdym = (w[CC(i, jp, km)] - w[CC(i, j, km)]) / dy;
                ma2 = 0.25 * (miu[CC(i, j, k)] + miu[CC(i, jp, k)] + miu[CC(ip, jp, k)] + miu[CC(ip, j, k)]);
                double ma5 = 0.25 * (miu[CC(i, j, k)] + miu[CC(i, jp, k)] + miu[CC(im, jp, k)] + miu[CC(im, j, k)]);
                double ma6 = 0.25 * (miu[CC(i, j, k)] + miu[CC(i, j, km)] + miu[CC(i, jp, km)] + miu[CC(i, jp, k)]);
                double ma7 = 0.25 * (miu[CC(i, j, k)] + miu[CC(i, j, kp)] + miu[CC(i, jp, kp)] + miu[CC(i, jp, k)]);
                RD1 = (dvdxp * ma2 - dvdxm * ma5) / dx;
                RD2 = (dudyp * ma2 - dudym * ma5) / dx;
                RD3 = 2. * (dvdyp * miu[CC(i, jp, k)] - dvdym * miu[CC(i, j, k)]) / dy;
                RD4 = (dvdzp * ma7 - dvdzm * ma6) / dz;
                RD5 = (dwdyp * ma7 - dwdym * ma6) / dz;
                RDv[c] = (RD1 + RD2 + RD3 + RD4 + RD5) / S->rho_v[c];

                dwdxp = (w[CC(ip, j, k)] - w[CC(i, j, k)]) / dx;
                dwdxm = (w[CC(i, j, k)] - w[CC(im, j, k)]) / dx;
                dudzp = (u[CC(i, j, kp)] - u[CC(i, j, k)]) / dz;
                dudzm = (u[CC(im, j, kp)] - u[CC(im, j, k)]) / dz;
                dwdyp = (w[CC(i, jp, k)] - w[CC(i, j, k)]) / dy;
                dwdym = (w[CC(i, j, k)] - w[CC(i, jm, k)]) / dy;
                dvdzp = (v[CC(i, j, kp)] - v[CC(i, j, k)]) / dz;
                dvdzm = (v[CC(i, jm, kp)] - v[CC(i, jm, k)]) / dz;
                double dwdzp = (w[CC(i, j, kp)] - w[CC(i, j, k)]) / dz;
                double dwdzm = (w[CC(i, j, k)] - w[CC(i, j, km)]) / dz;
                ma4 = 0.25 * (miu[CC(i, j, k)] + miu[CC(i, j, kp)] + miu[CC(ip, j, kp)] + miu[CC(ip, j, k)]);
                double ma8 = 0.25 * (miu[CC(i, j, k)] + miu[CC(i, j, kp)] + miu[CC(im, j, kp)] + miu[CC(im, j, k)]);
                ma7 = 0.25 * (miu[CC(i, j, k)] + miu[CC(i, j, kp)] + miu[CC(i, jp, kp)] + miu[CC(i, jp, k)]);
                double ma9 = 0.25 * (miu[CC(i, j, k)] + miu[CC(i, j, kp)] + miu[CC(i, jm, kp)] + miu[CC(i, jm, k)]);
                RD1 = (dwdxp * ma4 - dwdxm * ma8) / dx;
                RD2 = (dudzp * ma4 - dudzm * ma8) / dx;
                RD3 = (dwdyp * ma7 - dwdym * ma9) / dy;
                RD4 = (dvdzp * ma7 - dvdzm * ma9) / dy;
                RD5 = 2. * (dwdzp * miu[CC(i, j, kp)] - dwdzm * miu[CC(i, j, k)]) / dz;
                RDw[c] = (RD1 + RD2 + RD3 + RD4 + RD5) / S->rho_w[c];
            }
}

/* mom_main (mom.f90:13-62) with NS_grav (mom.f90:277-301) */
int oracle_mom_main(oracle_state* S, double* RHSu, double* RHSv, double* RHSw) {
    double *RCu = zalloc(S->ni), *RCv = zalloc(S->ni), *RCw = zalloc(S->ni);
    double *RDu = zalloc(S->ni), *RDv = zalloc(S->ni), *RDw = zalloc(S->ni);
    cntr2_conv(S, RCu, RCv, RCw);
    delta_diff(S, RDu, RDv, RDw);
    const double iFr = S->P.iFr, *g = S->P.g_unit;
    OMP_FOR
    for (long c = 0; c < S->ni; c++) {
        double RGu = iFr * (1. - S->rho_avr / S->rho_u[c]) * g[0];
        double RGv = iFr * (1. - S->rho_avr / S->rho_v[c]) * g[1];
        double RGw = iFr * (1. - S->rho_avr / S->rho_w[c]) * g[2];
        RHSu[c] = RCu[c] + RDu[c] * S->visc + RGu;
        RHSv[c] = RCv[c] + RDv[c] * S->visc + RGv;
        RHSw[c] = RCw[c] + RDw[c] * S->visc + RGw;
    }
    free(RCu); free(RCv); free(RCw); free(RDu); free(RDv); free(RDw);
    return 0;
}

/* Coefficients of ab2 for a step dt after a step dt_old, r = dt / dt_old:
 * c[0] RU^n - c[1] RU^(n-1) is the linear extrapolation of RU to t^n + dt/2 and
 * c[2] p^n - c[3] p^(n-1) that of p to t^n + dt.  At r = 1 they are 1.5, 0.5, 2 and 1,
 * the constants of ab2.f90:55-57 and 79-97; a variable dt is not in the reference. */
void oracle_ab2_coeffs(double dt, double dt_old, double c[4]) {
    const double r = dt_old > 0. ? dt / dt_old : 1.;
    c[0] = 1. + r / 2.;
    c[1] = r / 2.;
    c[2] = 1. + r;
    c[3] = r;
}

static double oracle_chkdt_speed(oracle_state* S);

/* dt from chkdt's speed for cfl > 0 (not in the reference, whose dt is fixed,
 * main.f90:216): dt = cfl / speed = 2 cfl dtmax, growing at most 1.2x per update */
static void oracle_adapt_dt(oracle_state* S, int limit_growth) {
    double dt = S->P.cfl / oracle_chkdt_speed(S);
    if (limit_growth && dt > 1.2 * S->dt) dt = 1.2 * S->dt;
    S->dt = dt;
}

/* ab2 with surface_tension_method = 'GFM' (ab2.f90:21-138), momxp/yp/zp (mom.f90:779-851) */
int oracle_ab2(oracle_state* S) {
    long ni = S->ni, nc = S->nc;
    double *dunew = zalloc(ni), *dvnew = zalloc(ni), *dwnew = zalloc(ni);
    double *phat = zalloc(nc), *pjx = zalloc(nc), *pjy = zalloc(nc), *pjz = zalloc(nc);
    const double dt = S->dt;
    double cf[4];
    oracle_ab2_coeffs(dt, S->dt_old, cf);
    const double abn = cf[0], abo = cf[1], exn = cf[2], exo = cf[3];
    S->dt_old = dt;
    oracle_mom_main(S, dunew, dvnew, dwnew);
    OMP_FOR
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) {
                long c = II(i, j, k), q = CC(i, j, k);
                S->dudt[q] = S->unew[q] + dt * (abn * dunew[c] - abo * S->duold[c]);
                S->dvdt[q] = S->vnew[q] + dt * (abn * dvnew[c] - abo * S->dvold[c]);
                S->dwdt[q] = S->wnew[q] + dt * (abn * dwnew[c] - abo * S->dwold[c]);
            }
    memcpy(S->duold, dunew, sizeof(double) * ni);
    memcpy(S->dvold, dvnew, sizeof(double) * ni);
    memcpy(S->dwold, dwnew, sizeof(double) * ni);
    OMP_FOR
    for (long q = 0; q < nc; q++) {
        pjx[q] = exn * S->p_x[q] - exo * S->p_xold[q];
        pjy[q] = exn * S->p_y[q] - exo * S->p_yold[q];
        pjz[q] = exn * S->p_z[q] - exo * S->p_zold[q];
        phat[q] = exn * S->pnew[q] - exo * S->pold[q];
    }
    const double rho0 = S->rho0;
    OMP_FOR
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) {
                long c = II(i, j, k);
                dunew[c] = -(1. / S->rho_u[c] - 1. / rho0) * S->dxi * (phat[CC(i + 1, j, k)] - pjx[CC(i, j, k)] - phat[CC(i, j, k)]);
                dvnew[c] = -(1. / S->rho_v[c] - 1. / rho0) * S->dyi * (phat[CC(i, j + 1, k)] - pjy[CC(i, j, k)] - phat[CC(i, j, k)]);
                dwnew[c] = -(1. / S->rho_w[c] - 1. / rho0) * S->dzi * (phat[CC(i, j, k + 1)] - pjz[CC(i, j, k)] - phat[CC(i, j, k)]);
            }
    double forceytot = S->P.constPoi ? -12. * S->visc / (S->lz * S->lz) : 0.;
    OMP_FOR
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) {
                long c = II(i, j, k), q = CC(i, j, k);
                S->dudt[q] = S->dudt[q] + dt * dunew[c];
                S->dvdt[q] = S->dvdt[q] + dt * dvnew[c] - forceytot * dt / S->rho_v[c];
                S->dwdt[q] = S->dwdt[q] + dt * dwnew[c];
            }
    bounduvw_arr(S, S->dudt, S->dvdt, S->dwdt);
    free(dunew); free(dvnew); free(dwnew); free(phat); free(pjx); free(pjy); free(pjz);
    return 0;
}

/* ---------------------------------------------------------------- projection */
/* fillps (fillps.f90:13-57) into the main-program array p */
int oracle_fillps(oracle_state* S) {
    const double coef = S->rho0 / S->dt;
    OMP_FOR
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) {
                double vel_x = (S->dudt[CC(i, j, k)] - S->dudt[CC(i - 1, j, k)]) * S->dxi;
                double vel_y = (S->dvdt[CC(i, j, k)] - S->dvdt[CC(i, j - 1, k)]) * S->dyi;
                double vel_z = (S->dwdt[CC(i, j, k)] - S->dwdt[CC(i, j, k - 1)]) * S->dzi;
                S->p[CC(i, j, k)] = coef * (vel_x + vel_y + vel_z);
            }
    return 0;
}

/* ipot (interface.f90:225-277) */
static void ipot(oracle_state* S) {
    int lmax = S->lmax;
    double* dtemp = zalloc((long)lmax * lmax);
    for (long q = 0; q < (long)lmax * lmax; q++) S->dmin[q] = 0.;
    const double cutoff = S->NB2, dz = S->dz;
    for (int m = 0; m < lmax; m++)
        for (int n = 0; n < lmax; n++) {
            if (n == m) continue;
            double d1 = 1e3;
            const double *pm = S->lvset[m], *pn = S->lvset[n];
            OMP_FOR_RED(min : d1)
            for (long q = 0; q < S->nls; q++) {
                double sp = 1e3;
                if (pm[q] > -1. * dz && pm[q] < 1. * dz && pn[q] < cutoff) sp = pm[q] + pn[q];
                if (sp < d1) d1 = sp;
            }
            dtemp[m + (long)lmax * n] = d1;
        }
    for (int m = 0; m < lmax - 1; m++)
        for (int n = m + 1; n < lmax; n++) S->dmin[m + (long)lmax * n] = 0.5 * (dtemp[m + (long)lmax * n] + dtemp[n + (long)lmax * m]);
    free(dtemp);
}

double oracle_ipot_dmin(oracle_state* S, int m, int n) { return S->dmin[m + (long)S->lmax * n]; }

/* gfm4p with NS_diff_discretization = 'del', surface_tension_method = 'GFM',
 * attractive = .false. (gfm.f90:17-267) */
int oracle_gfm4p(oracle_state* S) {
    const double nondi = 1. / S->P.We;
    const double dx2 = S->dx * S->dx, dy2 = S->dy * S->dy, dz2 = S->dz * S->dz;
    ipot(S);
    memcpy(S->p_xold, S->p_x, sizeof(double) * S->nc);
    memcpy(S->p_yold, S->p_y, sizeof(double) * S->nc);
    memcpy(S->p_zold, S->p_z, sizeof(double) * S->nc);
    OMP_FOR
    for (long q = 0; q < S->nc; q++) { S->p_jump[q] = 0.; S->p_x[q] = 0.; S->p_y[q] = 0.; S->p_z[q] = 0.; }
    double* re_curv = zalloc(S->nc);
    for (int l = 0; l < S->lmax; l++) {
        const double* ls = S->lvset[l];
        get_curvature_arr(S, ls, S->curv_cmn);
        boundc(S, S->curv_cmn);
        re_curvature_arr(S, ls, S->curv_cmn, re_curv);
        memcpy(S->curv_cmn, re_curv, sizeof(double) * S->nc);
        boundc(S, S->curv_cmn);
        const double* cvf = S->curv_cmn;
        OMP_FOR
        for (int k = 1; k <= S->kmax; k++)
            for (int j = 1; j <= S->jmax; j++)
                for (int i = 1; i <= S->imax; i++) {
                    if (!(fabs(ls[LS(i, j, k)]) <= 2. * S->dz)) continue;
                    double jp_x[3] = {0., 0., 0.}, jp_y[3] = {0., 0., 0.}, jp_z[3] = {0., 0., 0.};
                    double pxs = 0., pys = 0., pzs = 0., pjs;
                    double ph[3], cv[3];
                    for (int nx = -1; nx <= 1; nx += 2) {
                        for (int q = -1; q <= 1; q++) { ph[q + 1] = ls[LS(i + q, j, k)]; cv[q + 1] = cvf[CC(i + q, j, k)]; }
                        if (ph[1] * ph[1 + nx] < 0.) {
                            double icurv = oracle_icurvature(nx, ph, cv);
                            jp_x[nx + 1] = nondi * (-icurv);
                            if (nx == 1) pxs = jp_x[nx + 1];
                        }
                    }
                    double jump_x = jp_x[0] + jp_x[1] + jp_x[2];
                    for (int ny = -1; ny <= 1; ny += 2) {
                        for (int q = -1; q <= 1; q++) { ph[q + 1] = ls[LS(i, j + q, k)]; cv[q + 1] = cvf[CC(i, j + q, k)]; }
                        if (ph[1] * ph[1 + ny] < 0.) {
                            double icurv = oracle_icurvature(ny, ph, cv);
                            jp_y[ny + 1] = nondi * (-icurv);
                            if (ny == 1) pys = jp_y[ny + 1];
                        }
                    }
                    double jump_y = jp_y[0] + jp_y[1] + jp_y[2];
                    for (int nz = -1; nz <= 1; nz += 2) {
                        for (int q = -1; q <= 1; q++) { ph[q + 1] = ls[LS(i, j, k + q)]; cv[q + 1] = cvf[CC(i, j, k + q)]; }
                        if (ph[1] * ph[1 + nz] < 0.) {
                            double icurv = oracle_icurvature(nz, ph, cv);
                            jp_z[nz + 1] = nondi * (-icurv);
                            if (nz == 1) pzs = jp_z[nz + 1];
                        }
                    }
                    double jump_z = jp_z[0] + jp_z[1] + jp_z[2];
                    double phc = ls[LS(i, j, k)];
                    if (phc > 0.) {
                        pjs = (jump_x / dx2 + jump_y / dy2 + jump_z / dz2);
                    } else if (phc < 0.) {
                        pjs = -(jump_x / dx2 + jump_y / dy2 + jump_z / dz2);
                        pxs = -pxs; pys = -pys; pzs = -pzs;
                    } else {
                        pjs = 0.; pxs = 0.; pys = 0.; pzs = 0.;
                    }
                    long q = CC(i, j, k);
                    if (l == 0) { S->p_jump[q] = pjs; S->p_x[q] = pxs; S->p_y[q] = pys; S->p_z[q] = pzs; }
                    else {
                        S->p_jump[q] = S->p_jump[q] + pjs; S->p_x[q] = S->p_x[q] + pxs;
                        S->p_y[q] = S->p_y[q] + pys; S->p_z[q] = S->p_z[q] + pzs;
                    }
                }
    }
    free(re_curv);
    return 0;
}

int oracle_add_pjump(oracle_state* S) {
    OMP_FOR
    for (long q = 0; q < S->nc; q++) S->p[q] = S->p[q] + S->p_jump[q];
    return 0;
}

/* solver2d (solver.f90:68-196) on p(1:imax,1:jmax,1:kmax); the 2DECOMP transposes
 * are identities on one rank */
/* VCOSQB and VCOSQF (fft.f:218-370, 68-214) by the sums that define them
 * (fft.f:266-270, 116-120), not by the fast algorithm:
 *   VCOSQB: X(i) = sum_{k=1..n} 4 X(k) cos((2k-1)(i-1) pi/(2n)) / sqrt(4n)
 *   VCOSQF: X(i) = (X(1) + sum_{k=2..n} 2 X(k) cos((2i-1)(k-1) pi/(2n))) / sqrt(4n) */
void oracle_vcosqb(int n, double* x) {
    const double pi = acos(-1.), s = sqrt(4. * n);
    double* y = zalloc(n);
    for (int i = 1; i <= n; i++) {
        double acc = 0.;
        for (int k = 1; k <= n; k++) acc += 4. * x[k - 1] * cos((double)((2 * k - 1) * (i - 1)) * pi / (2. * n));
        y[i - 1] = acc / s;
    }
    memcpy(x, y, sizeof(double) * n);
    free(y);
}

void oracle_vcosqf(int n, double* x) {
    const double pi = acos(-1.), s = sqrt(4. * n);
    double* y = zalloc(n);
    for (int i = 1; i <= n; i++) {
        double acc = x[0];
        for (int k = 2; k <= n; k++) acc += 2. * x[k - 1] * cos((double)((2 * i - 1) * (k - 1)) * pi / (2. * n));
        y[i - 1] = acc / s;
    }
    memcpy(x, y, sizeof(double) * n);
    free(y);
}

/* solver2d_cos (solver.f90:202-317) on p's interior, with initsolver's branch for
 * BC_in_y = 'Walls' (1), 'InOut' (2), 'OutOut' (3) (initsolver.f90:84-160): forward real
 * FFT along x, VCOSQB along z, Gauss elimination along y with the eigenvalue
 * xrt(i) + zrt(k), VCOSQF along z, backward FFT along x. */
int oracle_solver2d_cos(oracle_state* S, int bcy) {
    const int imax = S->imax, jmax = S->jmax, kmax = S->kmax;
    const double pi = S->pi;
    int nl = imax > jmax ? imax : jmax;
    if (kmax > nl) nl = kmax;
    double* line = zalloc(nl);
    /* y tridiagonal (initsolver.f90:88-124) */
    double* a = zalloc(jmax);
    double* b = zalloc(jmax);
    double* c = zalloc(jmax);
    for (int j = 0; j < jmax; j++) { a[j] = S->dyi * S->dyi; c[j] = S->dyi * S->dyi; b[j] = -(a[j] + c[j]); }
    if (bcy == 1) { b[0] = b[0] + a[0]; b[jmax - 1] = b[jmax - 1] + c[jmax - 1]; }
    else if (bcy == 2) { b[0] = b[0] + a[0]; b[jmax - 1] = b[jmax - 1] - c[jmax - 1]; }
    else { b[0] = b[0] - a[0]; b[jmax - 1] = b[jmax - 1] - c[jmax - 1]; }
    a[0] = 0.; c[jmax - 1] = 0.;
    /* eigenvalues (initsolver.f90:132-153) */
    double* xrt = zalloc(imax + 1);
    for (int i = 3; i <= imax; i += 2) {
        double sn = sin((double)(i - 1) * pi / (2. * imax));
        xrt[i - 1] = -4. * S->dxi * S->dxi * (sn * sn);
        xrt[i] = xrt[i - 1];
    }
    xrt[1] = 0.; xrt[imax] = -4. * S->dxi * S->dxi;
    double* zrt = zalloc(kmax + 1);
    for (int k = 1; k <= kmax; k++) zrt[k] = (S->dzi * S->dzi) * (-2. + 2. * cos((double)(k - 1) * pi / kmax));
#define PZ(i, j, k) S->p[CC(i, j, k)]
    for (int k = 1; k <= kmax; k++)
        for (int j = 1; j <= jmax; j++) {
            for (int i = 1; i <= imax; i++) line[i - 1] = PZ(i, j, k);
            oracle_rfft_forward(imax, line);
            for (int i = 1; i <= imax; i++) PZ(i, j, k) = line[i - 1];
        }
    for (int j = 1; j <= jmax; j++)
        for (int i = 1; i <= imax; i++) {
            for (int k = 1; k <= kmax; k++) line[k - 1] = PZ(i, j, k);
            oracle_vcosqb(kmax, line);
            for (int k = 1; k <= kmax; k++) PZ(i, j, k) = line[k - 1];
        }
    /* Gauss elimination in y (solver.f90:253-290) */
    double* d = zalloc(jmax);
    for (int k = 1; k <= kmax; k++)
        for (int i = 1; i <= imax; i++) {
            const double xz = xrt[i] + zrt[k];
            double z = 1. / (b[0] + xz);
            d[0] = c[0] * z;
            PZ(i, 1, k) = PZ(i, 1, k) * z;
            for (int j = 2; j <= jmax - 1; j++) {
                double bb = b[j - 1] + xz;
                z = 1. / (bb - a[j - 1] * d[j - 2]);
                d[j - 1] = c[j - 1] * z;
                PZ(i, j, k) = (PZ(i, j, k) - a[j - 1] * PZ(i, j - 1, k)) * z;
            }
            double bb = b[jmax - 1] + xz;
            z = bb - a[jmax - 1] * d[jmax - 2];
            if (z != 0.) PZ(i, jmax, k) = (PZ(i, jmax, k) - a[jmax - 1] * PZ(i, jmax - 1, k)) / z;
            else PZ(i, jmax, k) = 0.;
            for (int j = jmax - 1; j >= 1; j--) PZ(i, j, k) = PZ(i, j, k) - d[j - 1] * PZ(i, j + 1, k);
        }
    for (int j = 1; j <= jmax; j++)
        for (int i = 1; i <= imax; i++) {
            for (int k = 1; k <= kmax; k++) line[k - 1] = PZ(i, j, k);
            oracle_vcosqf(kmax, line);
            for (int k = 1; k <= kmax; k++) PZ(i, j, k) = line[k - 1];
        }
    for (int k = 1; k <= kmax; k++)
        for (int j = 1; j <= jmax; j++) {
            for (int i = 1; i <= imax; i++) line[i - 1] = PZ(i, j, k);
            oracle_rfft_backward(imax, line);
            for (int i = 1; i <= imax; i++) PZ(i, j, k) = line[i - 1];
        }
#undef PZ
    free(line); free(a); free(b); free(c); free(xrt); free(zrt); free(d);
    return 0;
}

int oracle_solver2d(oracle_state* S) {
    int imax = S->imax, jmax = S->jmax, kmax = S->kmax;
    OMP_FOR
    for (int k = 1; k <= kmax; k++) {
        double* line = (double*)malloc(sizeof(double) * imax);
        for (int j = 1; j <= jmax; j++) {
            for (int i = 1; i <= imax; i++) line[i - 1] = S->p[CC(i, j, k)];
            oracle_rfft_forward(imax, line);
            for (int i = 1; i <= imax; i++) S->p[CC(i, j, k)] = line[i - 1];
        }
        free(line);
    }
    OMP_FOR
    for (int k = 1; k <= kmax; k++) {
        double* line = (double*)malloc(sizeof(double) * jmax);
        for (int i = 1; i <= imax; i++) {
            for (int j = 1; j <= jmax; j++) line[j - 1] = S->p[CC(i, j, k)];
            oracle_rfft_forward(jmax, line);
            for (int j = 1; j <= jmax; j++) S->p[CC(i, j, k)] = line[j - 1];
        }
        free(line);
    }
    double* d = zalloc((long)imax * jmax * kmax);
#define PZ(i, j, k) S->p[CC(i, j, k)]
#define DD(i, j, k) d[II(i, j, k)]
    const double *a = S->a, *b = S->b, *c = S->c;
    /* solver.f90:119-160, each (i, j) column through its own sweeps */
    OMP_FOR
    for (int j = 1; j <= jmax; j++)
        for (int i = 1; i <= imax; i++) {
            const double xy = S->xyrt[(i - 1) + (long)imax * (j - 1)];
            double z = 1. / (b[0] + xy);
            DD(i, j, 1) = c[0] * z;
            PZ(i, j, 1) = PZ(i, j, 1) * z;
            for (int k = 2; k <= kmax - 1; k++) {
                double bb = b[k - 1] + xy;
                z = 1. / (bb - a[k - 1] * DD(i, j, k - 1));
                DD(i, j, k) = c[k - 1] * z;
                PZ(i, j, k) = (PZ(i, j, k) - a[k - 1] * PZ(i, j, k - 1)) * z;
            }
            {
                double bb = b[kmax - 1] + xy;
                z = bb - a[kmax - 1] * DD(i, j, kmax - 1);
                if (z != 0.) PZ(i, j, kmax) = (PZ(i, j, kmax) - a[kmax - 1] * PZ(i, j, kmax - 1)) / z;
                else PZ(i, j, kmax) = 0.;
            }
            for (int k = kmax - 1; k >= 1; k--) PZ(i, j, k) = PZ(i, j, k) - DD(i, j, k) * PZ(i, j, k + 1);
        }
#undef PZ
#undef DD
    free(d);
    OMP_FOR
    for (int k = 1; k <= kmax; k++) {
        double* line = (double*)malloc(sizeof(double) * jmax);
        for (int i = 1; i <= imax; i++) {
            for (int j = 1; j <= jmax; j++) line[j - 1] = S->p[CC(i, j, k)];
            oracle_rfft_backward(jmax, line);
            for (int j = 1; j <= jmax; j++) S->p[CC(i, j, k)] = line[j - 1];
        }
        free(line);
    }
    OMP_FOR
    for (int k = 1; k <= kmax; k++) {
        double* line = (double*)malloc(sizeof(double) * imax);
        for (int j = 1; j <= jmax; j++) {
            for (int i = 1; i <= imax; i++) line[i - 1] = S->p[CC(i, j, k)];
            oracle_rfft_backward(imax, line);
            for (int i = 1; i <= imax; i++) S->p[CC(i, j, k)] = line[i - 1];
        }
        free(line);
    }
    return 0;
}

/* correc (correc.f90:15-44) */
int oracle_correc(oracle_state* S) {
    const double *p = S->p, dt = S->dt, rho0 = S->rho0;
    OMP_FOR
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) {
                long q = CC(i, j, k);
                S->unew[q] = S->dudt[q] - dt * S->dxi * (p[CC(i + 1, j, k)] - p[q] - S->p_x[q]) / rho0;
                S->vnew[q] = S->dvdt[q] - dt * S->dyi * (p[CC(i, j + 1, k)] - p[q] - S->p_y[q]) / rho0;
                S->wnew[q] = S->dwdt[q] - dt * S->dzi * (p[CC(i, j, k + 1)] - p[q] - S->p_z[q]) / rho0;
            }
    return 0;
}

static double* field_ptr(oracle_state* S, int field, int l) {
    switch (field) {
    case ICLS_F_UNEW: return S->unew;  case ICLS_F_VNEW: return S->vnew;  case ICLS_F_WNEW: return S->wnew;
    case ICLS_F_DUDT: return S->dudt;  case ICLS_F_DVDT: return S->dvdt;  case ICLS_F_DWDT: return S->dwdt;
    case ICLS_F_PNEW: return S->pnew;  case ICLS_F_POLD: return S->pold;  case ICLS_F_P: return S->p;
    case ICLS_F_PJUMP: return S->p_jump;
    case ICLS_F_DUOLD: return S->duold; case ICLS_F_DVOLD: return S->dvold; case ICLS_F_DWOLD: return S->dwold;
    case ICLS_F_PX: return S->p_x; case ICLS_F_PY: return S->p_y; case ICLS_F_PZ: return S->p_z;
    case ICLS_F_PXOLD: return S->p_xold; case ICLS_F_PYOLD: return S->p_yold; case ICLS_F_PZOLD: return S->p_zold;
    case ICLS_F_MIU: return S->miu; case ICLS_F_RHO: return S->rho; case ICLS_F_JIEMIAN: return S->jiemian;
    case ICLS_F_CURV: return S->curv_cmn;
    case ICLS_F_UCC: return S->uccnew; case ICLS_F_VCC: return S->vccnew; case ICLS_F_WCC: return S->wccnew;
    case ICLS_F_RHO_U: return S->rho_u; case ICLS_F_RHO_V: return S->rho_v; case ICLS_F_RHO_W: return S->rho_w;
    case ICLS_F_LVSET: return (l >= 0 && l < S->lmax) ? S->lvset[l] : NULL;
    case ICLS_F_NORMAL_M: return S->nrm_m;
    default: return NULL;
    }
}

static int field_shape(int field) {
    switch (field) {
    case ICLS_F_DUOLD: case ICLS_F_DVOLD: case ICLS_F_DWOLD: case ICLS_F_UCC: case ICLS_F_VCC: case ICLS_F_WCC:
    case ICLS_F_RHO_U: case ICLS_F_RHO_V: case ICLS_F_RHO_W: case ICLS_F_NORMAL_M: return ICLS_SHAPE_I;
    case ICLS_F_LVSET: return ICLS_SHAPE_LS;
    default: return ICLS_SHAPE_C;
    }
}

double* oracle_field(oracle_state* S, int field, int l, int* shape, long* count) {
    int sh = field_shape(field);
    if (shape) *shape = sh;
    if (count) *count = sh == ICLS_SHAPE_C ? S->nc : sh == ICLS_SHAPE_LS ? S->nls : S->ni;
    return field_ptr(S, field, l);
}

int oracle_boundp(oracle_state* S, int field) {
    double* p = field_ptr(S, field, 0);
    if (!p) return 1;
    boundp_arr(S, p);
    return 0;
}

int oracle_bounduvw(oracle_state* S, int which) {
    if (which == 0) bounduvw_arr(S, S->unew, S->vnew, S->wnew);
    else bounduvw_arr(S, S->dudt, S->dvdt, S->dwdt);
    return 0;
}

/* main.f90:164-185 */
int oracle_update_pressure(oracle_state* S, int first_step) {
    memcpy(S->pold, S->pnew, sizeof(double) * S->nc);
    memcpy(S->pnew, S->p, sizeof(double) * S->nc);
    boundp_arr(S, S->pnew);
    const int k = 1;
    double norm = 0.;
    for (int i = 0; i <= S->i1; i++) {
        double s = 0.;
        for (int j = 0; j <= S->j1; j++) s = s + S->pnew[CC(i, j, k)];
        norm = norm + s;
    }
    norm = norm / (1. * S->imax * S->jmax);
    OMP_FOR
    for (long q = 0; q < S->nc; q++) S->pnew[q] = S->pnew[q] - norm;
    const double norm_old = S->norm_old;
    OMP_FOR
    for (long q = 0; q < S->nc; q++) S->pold[q] = first_step ? S->pold[q] - norm : S->pold[q] + norm_old - norm;
    S->norm_old = norm;
    return 0;
}

/* average_vel_at_center (misc.f90:587-604) */
int oracle_average_vel_at_center(oracle_state* S) {
    OMP_FOR
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) {
                long c = II(i, j, k);
                S->uccnew[c] = (S->unew[CC(i - 1, j, k)] + S->unew[CC(i, j, k)]) / 2.;
                S->vccnew[c] = (S->vnew[CC(i, j - 1, k)] + S->vnew[CC(i, j, k)]) / 2.;
                S->wccnew[c] = (S->wnew[CC(i, j, k - 1)] + S->wnew[CC(i, j, k)]) / 2.;
            }
    return 0;
}

/* average_stream_vel (misc.f90:642-655) */
double oracle_average_stream_vel(oracle_state* S) {
    double s = 0.;
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) s = s + S->vnew[CC(i, j, k)];
    s = s / (S->imax * S->jmax * S->kmax);
    S->v_bulk = s;
    return s;
}

/* chkdiv (chkdiv.f90:10-90); returns 1 if the 1e-12 abort would fire */
int oracle_chkdiv(oracle_state* S, int istep, double* divtot_o, double* divmax_o, int* ijk) {
    double divmax = 0., divtot = 0.;
    int im = 0, jm = 0, km = 0;
    for (int k = 1; k <= S->kmax; k++)
        for (int j = 1; j <= S->jmax; j++)
            for (int i = 1; i <= S->imax; i++) {
                double div = (S->wnew[CC(i, j, k)] - S->wnew[CC(i, j, k - 1)]) * S->dzi +
                             (S->vnew[CC(i, j, k)] - S->vnew[CC(i, j - 1, k)]) * S->dyi +
                             (S->unew[CC(i, j, k)] - S->unew[CC(i - 1, j, k)]) * S->dxi;
                divtot = divtot + div;
                div = fabs(div);
                if (div > divmax) { divmax = div; im = i; jm = j; km = k; }
            }
    if (divtot_o) *divtot_o = divtot;
    if (divmax_o) *divmax_o = divmax;
    if (ijk) { ijk[0] = im; ijk[1] = jm; ijk[2] = km; }
    return (divmax > 1e-12 && istep != 0) ? 1 : 0;
}

/* chkdt's combined speed (chkdt.f90:11-104) */
static double oracle_chkdt_speed(oracle_state* S) {
    double u_max = 0., v_max = 0., w_max = 0.;
    for (int k = 1; k <= S->k1; k++)
        for (int j = 1; j <= S->j1; j++)
            for (int i = 1; i <= S->i1; i++) {
                double u = fabs((S->unew[CC(i, j, k)] + S->unew[CC(i - 1, j, k)]) / 2.);
                double v = fabs((S->vnew[CC(i, j, k)] + S->vnew[CC(i, j - 1, k)]) / 2.);
                double w = fabs((S->wnew[CC(i, j, k)] + S->wnew[CC(i, j, k - 1)]) / 2.);
                if (u > u_max) u_max = u;
                if (v > v_max) v_max = v;
                if (w > w_max) w_max = w;
            }
    double speed_conv = u_max / S->dx + v_max / S->dy + w_max / S->dz;
    const double *g = S->P.g_unit, iFr = S->P.iFr, ra = S->rho_avr, r1 = S->P.rho1, r2 = S->P.rho2;
    double gs[6] = {fabs(iFr * (1. - ra / r1) * g[0]), fabs(iFr * (1. - ra / r2) * g[0]),
                    fabs(iFr * (1. - ra / r1) * g[1]), fabs(iFr * (1. - ra / r2) * g[1]),
                    fabs(iFr * (1. - ra / r1) * g[2]), fabs(iFr * (1. - ra / r2) * g[2])};
    double g_max = gs[0];
    for (int q = 1; q < 6; q++) if (gs[q] > g_max) g_max = gs[q];
    double h_min = fmin(S->dx, fmin(S->dy, S->dz));
    double speed_grav = sqrt(g_max / h_min);
    double v1 = fmax(S->P.miu1 / S->P.rho1, S->P.miu2 / S->P.rho2) * S->visc;
    double v2 = 2. / (S->dx * S->dx) + 2. / (S->dy * S->dy) + 2. / (S->dz * S->dz);
    double speed_visc = v1 * v2;
    double curv_max = 2. / h_min;
    double speed_surf = sqrt(curv_max / (S->P.We * S->rho0 * (h_min * h_min)));
    double s1 = speed_conv + speed_visc;
    return (s1 + sqrt(s1 * s1 + 4. * (speed_grav * speed_grav) + 4. * (speed_surf * speed_surf))) / 2.;
}

/* chkdt (chkdt.f90:11-105) */
double oracle_chkdt(oracle_state* S) { return 0.5 / oracle_chkdt_speed(S); }

/* max_vel_at_center (misc.f90:610-636) */
double oracle_max_vel_at_center(oracle_state* S) {
    double m = 0.;
    for (long c = 0; c < S->ni; c++) {
        double v = sqrt(S->uccnew[c] * S->uccnew[c] + S->vccnew[c] * S->vccnew[c] + S->wccnew[c] * S->wccnew[c]);
        if (c == 0 || v > m) m = v;
    }
    return m;
}

/* ---------------------------------------------------------------- init_program (init.f90:22-117) */
int oracle_init_derived(oracle_state* S) {
    oracle_pseudo_inverse(S->dx, S->dy, S->dz, S->pinvA);  /* init_matrix */
    initsolver(S);
    S->dt = S->P.timestep;
    bounduvw_arr(S, S->unew, S->vnew, S->wnew);
    bounduvw_arr(S, S->dudt, S->dvdt, S->dwdt);
    oracle_chkdiv(S, 0, NULL, NULL, NULL);
    boundmls(S);
    boundp_arr(S, S->pnew);
    oracle_average_vel_at_center(S);
    oracle_update_materials(S);
    if (S->P.cfl > 0.) oracle_adapt_dt(S, 0);
    S->dt_old = S->dt;
    for (int l = 0; l < S->lmax; l++) S->init_volume[l] = S->volume[l];
    oracle_mom_main(S, S->duold, S->dvold, S->dwold);
    memcpy(S->pold, S->pnew, sizeof(double) * S->nc);
    memcpy(S->p_xold, S->p_x, sizeof(double) * S->nc);
    memcpy(S->p_yold, S->p_y, sizeof(double) * S->nc);
    memcpy(S->p_zold, S->p_z, sizeof(double) * S->nc);
    return 0;
}

int oracle_init_program(oracle_state* S) {
    S->time = 0.;
    init_vel_p(S);
    init_levelset(S);
    return oracle_init_derived(S);
}

/* one pass of the main loop (main.f90:105-210); returns ICLS_ERR_NAN on NaN */
int oracle_step(oracle_state* S, int istep, int begin) {
    S->time = S->time + S->dt;
    oracle_ls_main(S, istep);
    oracle_average_stream_vel(S);
    oracle_ab2(S);
    oracle_fillps(S);
    oracle_gfm4p(S);
    oracle_add_pjump(S);
    oracle_solver2d(S);
    boundp_arr(S, S->p);
    oracle_correc(S);
    bounduvw_arr(S, S->unew, S->vnew, S->wnew);
    oracle_update_pressure(S, istep == begin + 1);
    oracle_average_vel_at_center(S);
    if (isnan(S->vnew[CC(1, 1, 1)])) return ICLS_ERR_NAN;
    if (S->P.cfl > 0. && S->P.ioutchk > 0 && istep % S->P.ioutchk == 0) oracle_adapt_dt(S, 1);
    return 0;
}

double oracle_scalar(oracle_state* S, int which) {
    switch (which) {
    case ICLS_S_TIME: return S->time;
    case ICLS_S_DT: return S->dt;
    case ICLS_S_RHO_AVR: return S->rho_avr;
    case ICLS_S_V_BULK: return S->v_bulk;
    case ICLS_S_MAXCHK: return S->maxchk_glb;
    case ICLS_S_REINIT_RES: return S->reinit_residual;
    case ICLS_S_NORM_OLD: return S->norm_old;
    default: return NAN;
    }
}

void oracle_set_scalar(oracle_state* S, int which, double v) {
    switch (which) {
    case ICLS_S_TIME: S->time = v; break;
    case ICLS_S_DT: S->dt = v; break;
    case ICLS_S_RHO_AVR: S->rho_avr = v; break;
    case ICLS_S_NORM_OLD: S->norm_old = v; break;
    default: break;
    }
}

void oracle_set_volume(oracle_state* S, int l, double volume, double init_volume) {
    S->volume[l] = volume;
    S->init_volume[l] = init_volume;
}

double oracle_volume(oracle_state* S, int l, double* init_volume) {
    if (init_volume) *init_volume = S->init_volume[l];
    return S->volume[l];
}

const double* oracle_pseudo_inv_A(oracle_state* S) { return S->pinvA; }
