/*
 * icls.h -- C ABI of the MI355X ICLS step library (libicls.so).
 *
 * Every entry point replaces one Fortran subroutine of the reference's per-timestep
 * two-phase step; the citation on each line names the subroutine it stands in for
 * (paths relative to the reference tree).  Fortran host code binds these through
 * iso_c_binding (see the modules under icls-release_amd/fortran/).
 *
 * Conventions
 *   - Plain C types only: int32, double, pointers and sizes.  No torch types.
 *   - Every entry returns int32: ICLS_OK (0) or one of the ICLS_ERR_* codes below.
 *     icls_last_error() returns a human-readable message for the last failure.
 *   - One process-wide context (the reference keeps its state in module globals,
 *     src/common.f90:1-88).  Calls are not re-entrant and run on one HIP stream.
 *   - Host arrays passed to icls_upload / icls_download use the reference's own
 *     Fortran layout (column-major, i fastest) including its halo bounds:
 *       ICLS_SHAPE_C   (0:i1,0:j1,0:k1)            src/common.f90:7-9
 *       ICLS_SHAPE_LS  (-2:i1+2,-2:j1+2,-2:k1+2)   src/common.f90:17-18
 *       ICLS_SHAPE_I   (1:imax,1:jmax,1:kmax)      src/common.f90:11-13
 *     On the device x and y are periodic by index wrap (no halo storage); download
 *     fills the x/y halos of the host array from the periodic images.
 */
#ifndef ICLS_H
#define ICLS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define ICLS_OK              0
#define ICLS_ERR_PARAM       1   /* bad parameter block or argument                  */
#define ICLS_ERR_HIP         2   /* HIP runtime error (allocation, launch, copy)      */
#define ICLS_ERR_NAN         3   /* NaN in vnew(1,1,1): src/main.f90:207-210           */
#define ICLS_ERR_DIVERGENCE  4   /* max |div u| > 1e-12: src/chkdiv.f90:60-86          */
#define ICLS_ERR_STATE       5   /* call out of order (no context, not initialised)    */
#define ICLS_ERR_NODEVICE    6   /* no GPU visible                                     */
#define ICLS_ERR_UNSUPPORTED 7   /* option the reference aborts on or this library omits */

/* ---- precision modes ---- */
/* BC_in_y of the cosine solver (initsolver.f90:84-124) */
#define ICLS_BCY_WALLS          1
#define ICLS_BCY_INOUT          2
#define ICLS_BCY_OUTOUT         3
#define ICLS_MODE_FP64          0   /* every field and kernel in FP64 (reference precision)   */
#define ICLS_MODE_MIXED         1   /* FP32 level set, momentum, FFT; FP64 Thomas, residual,
                                       pressure accumulation and volume sums                 */
#define ICLS_MODE_MIXED_LEARNED 2   /* MIXED + MLP curvature closure on under-resolved cells  */

/* ---- level-set advection schemes (src/param.f90:216, mod_param_ls) ---- */
#define ICLS_ADV_HOUC5 0
#define ICLS_ADV_WENO5 1

/* ---- initial conditions (src/init.f90:146-210, 717-755, 901-986) ---- */
#define ICLS_INIU_ZER 0     /* 'zer': fluid at rest        */
#define ICLS_INIU_COU 1     /* 'cou': plane Couette profile */
#define ICLS_INIL_DRP 0     /* 'drp'/'mdp': one level set per drop (sphere SDF) */

#define ICLS_MAX_DROPS 1024

/*
 * Runtime parameter block.  Replaces the compile-time PARAMETERs of
 * src/param.f90:12-172 and src/param.f90:199-275 (mod_param_ls).
 * Grid spacing follows the reference's derivation: dxi = itot/lx, dx = 1/dxi
 * (src/param.f90:37-38).
 */
typedef struct icls_params {
    int32_t itot, jtot, ktot;        /* grid (param.f90:27)                              */
    int32_t lmax;                    /* number of level sets = drops (param.f90:170)     */
    double  lx, ly, lz;              /* domain size (param.f90:33-35)                    */
    double  timestep;                /* dt (param.f90:47)                                */
    double  rho1, rho2, miu1, miu2;  /* phase 1 = phi>0 (outer), phase 2 = drop          */
    double  Re, We, iFr;             /* param.f90:69-71                                  */
    double  g_unit[3];               /* param.f90:77                                     */
    int32_t ls_adv_space;            /* ICLS_ADV_HOUC5 / ICLS_ADV_WENO5                  */
    int32_t re_initialization;       /* param.f90:164                                    */
    int32_t reinit_step;             /* param.f90:230                                    */
    int32_t reinit_iter_max;         /* param.f90:233                                    */
    int32_t mass_correction;         /* param.f90:165.  1: the reference's inflation
                                        (levelset.f90:377-582), which restores about
                                        1/(2 dz) times the volume loss it targets and so
                                        overshoots and oscillates once it fires; 2: the
                                        same correction normalised by sum(delta^2 kappa
                                        |grad phi|^2) dV, which restores the loss itself
                                        (an extension); 0: off                            */
    int32_t inflate_step;            /* param.f90:258                                    */
    int32_t ns_np;                   /* no-slip walls at rest (param.f90:95)             */
    double  wall_speed;              /* |v| of the z walls when ns_np == 0; the reference
                                        hard-codes 1.5 (bound.f90:33-34, init.f90:193-194) */
    int32_t constPoi;                /* constant pressure gradient forcing (param.f90:96) */
    int32_t iniu;                    /* ICLS_INIU_*                                      */
    int32_t mode;                    /* ICLS_MODE_*                                      */
    int32_t ioutchk;                 /* diagnostics cadence (param.f90:85)               */
    double  cfl;                     /* 0: dt = timestep for the whole run, as the reference
                                        (main.f90:216).  > 0: dt = cfl / speed, with chkdt's
                                        combined speed (chkdt.f90:62-104), set at init and
                                        after every ioutchk-th step, growing at most 1.2x per
                                        update; AB2 and the split-pressure extrapolation then
                                        take their variable-step coefficients (icls_step).
                                        The reference's CFL is timestep/dz (param.f90:47) and
                                        its dt never adapts: this is an extension.          */
    int32_t ls_full;                 /* with lmax > 1 each level set is stepped inside a
                                        window around its drop (band + 9 cells, refreshed
                                        every 4 steps; the level-set arrays stay full-grid,
                                        values beyond the band outside the window are not
                                        updated); 1 = every level set over the whole grid */
    int32_t mass_trigger_exp;        /* the mass correction fires when |V/V0 - 1| > 1e-5
                                        (levelset.f90:392); n > 0 lowers the trigger to
                                        10^-n (an extension: the reference's 1e-5 leaves
                                        the volume free to drift by that much)             */
    int32_t reserved[4];
    double  drop_xc[ICLS_MAX_DROPS]; /* drop centres and radii, one level set each       */
    double  drop_yc[ICLS_MAX_DROPS];
    double  drop_zc[ICLS_MAX_DROPS];
    double  drop_r[ICLS_MAX_DROPS];
} icls_params;

/* ---- field identifiers for upload / download ---- */
enum icls_field {
    ICLS_F_UNEW = 0, ICLS_F_VNEW, ICLS_F_WNEW,     /* C  velocity          common.f90:7   */
    ICLS_F_DUDT, ICLS_F_DVDT, ICLS_F_DWDT,         /* C  predictor u*      common.f90:7   */
    ICLS_F_PNEW, ICLS_F_POLD, ICLS_F_P,            /* C  pressures         main.f90:57,60 */
    ICLS_F_PJUMP,                                  /* C  GFM jump          main.f90:57    */
    ICLS_F_DUOLD, ICLS_F_DVOLD, ICLS_F_DWOLD,      /* I  AB2 history       main.f90:56    */
    ICLS_F_PX, ICLS_F_PY, ICLS_F_PZ,               /* C  face jumps        common.f90:42  */
    ICLS_F_PXOLD, ICLS_F_PYOLD, ICLS_F_PZOLD,      /* C                    common.f90:43  */
    ICLS_F_MIU, ICLS_F_RHO, ICLS_F_JIEMIAN,        /* C  materials         common.f90:9   */
    ICLS_F_CURV,                                   /* C  curv_cmn          common.f90:19  */
    ICLS_F_UCC, ICLS_F_VCC, ICLS_F_WCC,            /* I  cell-centre vel.  common.f90:11  */
    ICLS_F_RHO_U, ICLS_F_RHO_V, ICLS_F_RHO_W,      /* I  face densities    common.f90:13  */
    ICLS_F_LVSET,                                  /* LS level set l       common.f90:18  */
    ICLS_F_NORMAL_M,                               /* I  |grad phi|        common.f90:25  */
    ICLS_F_COUNT
};

#define ICLS_SHAPE_C  0
#define ICLS_SHAPE_LS 1
#define ICLS_SHAPE_I  2

/* ---- scalar identifiers ---- */
enum icls_scalar {
    ICLS_S_TIME = 0,       /* time                     common.f90:71   */
    ICLS_S_DT,             /* dt                       common.f90:71   */
    ICLS_S_RHO_AVR,        /* rho_avr                  param.f90:62    */
    ICLS_S_V_BULK,         /* v_bulk                   common.f90:79   */
    ICLS_S_MAXCHK,         /* maxchk_glb               common.f90:48   */
    ICLS_S_REINIT_RES,     /* last re-init residual    levelset.f90:831 */
    ICLS_S_NORM_OLD,       /* pressure norm            main.f90:183    */
    ICLS_S_COUNT
};

/* ---- lifecycle ---- */
const char* icls_last_error(void);
int32_t icls_version(void);
/* Allocate the device state for this parameter block (replaces module allocation,
 * initmpi.f90 + common.f90).  Fails with ICLS_ERR_NODEVICE without a GPU. */
int32_t icls_init(const icls_params* p);
int32_t icls_finalize(void);
int32_t icls_field_shape(int32_t field, int32_t* shape, int64_t* count);

/* init_program (src/init.f90:22-117) for begin == 0: initial fields, least-squares
 * matrix, solver tables, boundary conditions, materials, init_volume, AB2 start. */
int32_t icls_init_program(void);
/* The part of init_program after the initial fields (init.f90:63-112), run on the
 * fields currently on the device (after icls_upload). */
int32_t icls_init_derived(void);

/* ---- field / scalar transfer ---- */
int32_t icls_upload(int32_t field, int32_t l, const double* host);
int32_t icls_download(int32_t field, int32_t l, double* host);
int32_t icls_get_scalar(int32_t which, double* value);
int32_t icls_set_scalar(int32_t which, double value);
int32_t icls_get_volume(int32_t l, double* volume, double* init_volume);
/* Restore volume(l) and init_volume(l) (common.f90:52), e.g. on restart. */
int32_t icls_set_volume(int32_t l, double volume, double init_volume);
int32_t icls_sync(void);

/* ---- the per-step operators (src/main.f90:103-291) ---- */
int32_t icls_ls_main(int32_t istep);                     /* levelset.f90:28-76          */
int32_t icls_average_stream_vel(double* v_avr);          /* misc.f90:642-655            */
int32_t icls_ab2(void);                                  /* ab2.f90:21-138              */
int32_t icls_fillps(void);                               /* fillps.f90:13-57 (into p)   */
int32_t icls_gfm4p(void);                                /* gfm.f90:17-267 (p_jump)     */
int32_t icls_add_pjump(void);                            /* main.f90:142  p = p+p_jump  */
int32_t icls_solver2d(void);                             /* solver.f90:68-196 (on p)    */
/* solver2d_cos (solver.f90:202-317) on p: FFT along x, VCOSQB along z, Gauss
 * elimination along y with the y boundary rule bc_y (initsolver.f90:84-124, BC_in_y:
 * ICLS_BCY_WALLS, ICLS_BCY_INOUT, ICLS_BCY_OUTOUT), VCOSQF along z, inverse FFT.
 * Computed in FP64 in every precision mode, with the library's own scratch field as
 * work space.  The other operators of this library keep y periodic: the entry serves a
 * caller that builds its own y-wall step around it. */
int32_t icls_solver2d_cos(int32_t bc_y);
int32_t icls_boundp(int32_t field);                      /* bound.f90:182-287           */
int32_t icls_correc(void);                               /* correc.f90:15-44            */
int32_t icls_bounduvw(int32_t which);                    /* bound.f90:15-177; 0=unew,1=dudt */
int32_t icls_update_pressure(int32_t first_step);        /* main.f90:164-185            */
int32_t icls_average_vel_at_center(void);                /* misc.f90:587-604            */
int32_t icls_check_nan(void);                            /* main.f90:207-210            */
int32_t icls_chkdiv(int32_t istep, double* divtot, double* divmax, int32_t* ijk); /* chkdiv.f90:10-90 */
/* No counterpart in the reference.  The divergence check of icls_chkdiv over caller-owned
 * device arrays, for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): it allocates nothing
 * and needs no context.  u, v, w: nx * ny * (nz + 6) elements of elem_bytes (8 double,
 * 4 float) in the library's layout, element (i0, j0, k) at i0 + nx (j0 + ny (k + 2)) for
 * k = -2 .. nz + 3 (x and y periodic, no halo; z ghost layer k = 0 read).  out: device
 * memory for 3 doubles, written on `stream` (a hipStream_t, NULL for the null stream):
 * divtot, max |div| and the linear index i0 + nx (j0 + ny (k - 1)) of the first cell
 * holding it.  The block partials live in one fixed 1.5 MiB device array per GPU, held
 * until the process ends; grids needing more than 65536 reduction blocks (beyond 512^3)
 * are rejected with ICLS_ERR_PARAM.  Calls that overlap on one GPU share that array and
 * race.  Returns after queueing; synchronise `stream` before reading out. */
int32_t icls_chkdiv_noalloc(const void* u, const void* v, const void* w, int32_t elem_bytes, int32_t nx, int32_t ny,
                            int32_t nz, double lx, double ly, double lz, double* out, void* stream);
int32_t icls_chkdt(double* dtmax);                       /* chkdt.f90:11-105            */
int32_t icls_max_vel_at_center(double* vmax);            /* misc.f90:610-636            */
/* spurious_currents (misc.f90:802-843): the static-drop capillary number
 * Ca = mu max|u_c| / sigma with the reference's fixed mu = 0.1, sigma = 1. */
int32_t icls_spurious_currents(double* Ca);
/* ipot (interface.f90:225-277): minimal distances between interfaces, dmin(lmax,lmax)
 * column-major, on request.  In the reference only exjump (the 'attractive' option,
 * not built) and the per-output diagnostic read dmin, so it is not part of icls_step. */
int32_t icls_ipot(double* dmin);

/* ---- checkpoint / restart ----
 * loadflds (src/loadd.f90:16-108) writes the file fld####### the reference reads back:
 * unew, vnew, wnew, pnew, lvset(l) for l = 1..lmax and jiemian, each the interior
 * (1:imax,1:jmax,1:kmax) in FP64, x fastest, native byte order; then time, nr, dt. */
int32_t icls_write_fld(const char* path, int32_t nr);
/* init_program's restart branch (src/init.f90:44-51, 63-110): init_vel_p, the fields of
 * the file, then the derived state as icls_init_derived (which resets dt to timestep
 * and init_volume to the restarted volume, as the reference does).  *begin receives nr. */
int32_t icls_init_program_restart(const char* path, int32_t* begin);

/* ---- output files ----
 * write_vtk / vtk (src/write_vtk.f90:10-39, 214-432): legacy binary VTK, big-endian FP64
 * STRUCTURED_POINTS at the cell centres (DIMENSIONS itot jtot ktot, ORIGIN dx/2.., SPACING
 * dx..) with the scalars U, V, W (cell-centre velocities), abs(Vel), V_r (v minus the
 * laminar profile 6 z (1 - z)), Pressure (pnew) and Interface (jiemian).  The reference
 * names the file data/para<istep:7>.vtk; the Fortran wrapper write_vtk(istep) does too. */
int32_t icls_write_vtk(const char* path);
/* post2d (src/output.f90:179-427): Tecplot ASCII of the mid-yz plane i = itot/2, nine
 * columns (9E15.7): y, z, u, v, w, p, level set 1, jiemian, fixhv (0: no static level set).
 * The reference names the file data/yz_mid<istep:7>.dat. */
int32_t icls_post2d(const char* path);
/* One whole time step: main.f90:105-210 in the reference order, with the kernels
 * fused where the arithmetic per element is unchanged.  AB2 (ab2.f90:55-57) and the
 * pressure extrapolation p_hat (ab2.f90:79-97) use the variable-step forms with
 * r = dt / dt_old, dt_old the dt of the previous ab2 (icls_ab2 or icls_step; at
 * init, dt itself):  u* = u + dt ((1 + r/2) RU^n - (r/2) RU^(n-1)),
 * p_hat = (1 + r) p^n - r p^(n-1); at r = 1 these are the reference's 1.5, 0.5, 2, 1
 * exactly.  With cfl > 0 and istep a multiple of ioutchk, the step ends by setting
 * the next step's dt from chkdt (one host synchronisation). */
int32_t icls_step(int32_t istep, int32_t begin);

/* ---- instrumentation ---- */
/* Per-kernel HIP-event timing on the library's stream.  When enabled every launch
 * is bracketed with events; icls_kernel_times copies up to n (name, total ms, count). */
int32_t icls_profile_enable(int32_t on);
/* Restrict event bracketing to launches named `name` (NULL or "" = all). */
int32_t icls_profile_only(const char* name);
int32_t icls_profile_reset(void);
int32_t icls_kernel_times(int32_t n, char* names /* n*64 */, double* total_ms, int64_t* counts, int32_t* n_out);

/* ---- learned curvature closure ----
 * No counterpart in the reference: an MLP (30 -> 64 -> 64 -> 1, tanh, BF16 matrix
 * cores with FP32 accumulation) that, in ICLS_MODE_MIXED_LEARNED, replaces the
 * smoothed curvature of get_curvature + re_curvature (src/interface.f90:20-119,
 * 186-218) on under-resolved interface cells -- |phi| < 2 dz, |grad phi| >= 0.9 and
 * |kappa| dz >= gate (default 0.25) -- with the cell's own least-squares curvature plus
 * the network's correction.  The smoothed finite-difference value is kept where the
 * correction exceeds tol / dz (default 0.5).
 *   input  (32): the 3x3x3 phi stencil (i fastest, then j, then k) less its centre
 *                value and least-squares slope, over dz, with the centre value / dz
 *                at position 13; the unit normal of the least-squares gradient; two
 *                zeros
 *   output     : the correction to the least-squares kappa * dz
 *   weights    : ICLS_MLP_NPARAM floats, W1 (64 x 32, row-major), b1 (64),
 *                W2 (64 x 64), b2 (64), w3 (64), b3 (1)
 * Training runs on the device: samples are spheres and their complements of radius
 * r_min .. r_max dz whose reference curvature is the solver's own operator on a 4x
 * finer grid; the target is its difference from the coarse operator (times dz), the
 * loss the mean squared error, minimised with Adam (cosine learning-rate decay). */
#define ICLS_MLP_NPARAM 6337
typedef struct {
    int64_t n_samples;    /* training samples (default 131072)                 */
    int32_t epochs;       /* passes over them (default 60)                     */
    int32_t batch;        /* rows per Adam step, a multiple of 16 (1024)       */
    float lr;             /* initial Adam learning rate (default 3e-3)         */
    float r_min, r_max;   /* sample radii in dz (default 1.5, 12)              */
    uint64_t seed;        /* sample generator seed                             */
    float loss_first;     /* out: mean loss of the first epoch                 */
    float loss_last;      /* out: mean loss of the last epoch                  */
} icls_mlp_train_opts;
int32_t icls_mlp_defaults(icls_mlp_train_opts* o);
int32_t icls_mlp_train(icls_mlp_train_opts* o);
int32_t icls_mlp_set_weights(const float* w, int32_t n);
int32_t icls_mlp_get_weights(float* w, int32_t n);
int32_t icls_mlp_set_gate(float gate, float tol);
/* Host-buffer entries for checks: n rows of 32 bf16 features (n a multiple of 16). */
int32_t icls_mlp_forward(const uint16_t* x_bf16, int32_t n, float* y);
/* Gradient of the mean squared error: ICLS_MLP_NPARAM values, then the loss. */
int32_t icls_mlp_gradient(const uint16_t* x_bf16, const float* t, int32_t n, float* grad_loss);
/* The training generator's samples: features, targets (the correction times dz) and,
 * if fd is not NULL, the coarse least-squares kappa * dz. */
int32_t icls_mlp_samples(int64_t n, uint64_t seed, float r_min, float r_max, uint16_t* x_bf16, float* t, float* fd);
/* Cells the closure was consulted on in the last gfm4p pass, summed over the level sets
 * (each level set's list capped at the list size). */
int32_t icls_mlp_listed(int32_t* n);

#ifdef __cplusplus
}
#endif
#endif /* ICLS_H */
