// icls_internal.hpp -- shared definitions of the MI355X ICLS step library.
//
// Device data layout (every 3D field, whatever its Fortran bounds):
//   x and y are stored without halos and are periodic by index wrap;
//   z carries three ghost layers on each side, k = -2 .. ktot+3 (Fortran numbering),
//   stored at k_store = k + 2.  Element (i0, j0, k) with 0-based i0, j0 sits at
//       i0 + nx * (j0 + ny * (k + 2)).
//   Rows along x are contiguous and 256-byte aligned for nx a multiple of 32 (FP64).
// Fortran C-halo arrays (0:i1,0:j1,0:k1) use k = 0 .. ktot+1 of this layout, LS-halo
// arrays (-2:i1+2,...) use all of it, interior arrays (1:imax,...) use k = 1 .. ktot.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/icls.h"

namespace icls {

struct Grid {
    int nx, ny, nz;
    long sy;   // stride of j (= nx)
    long sz;   // stride of k (= nx*ny)
    double dx, dy, dz, dxi, dyi, dzi;
    int pow2;  // dxi, dyi, dzi are powers of two: a/dx == a*dxi exactly
};

__host__ __device__ inline long gidx(const Grid& g, int i0, int j0, int k) {
    return (long)i0 + g.sy * (long)j0 + g.sz * (long)(k + 2);
}
__host__ __device__ inline int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Scalar slots in the device scalar block (FP64).
enum DevScalar {
    DS_RHO_AVR = 0, DS_V_BULK, DS_NORM, DS_NORM_OLD, DS_REINIT_RES, DS_MAXCHK, DS_AREA, DS_INFL_TEMP,
    DS_INFL_FLAG, DS_DIV_TOT, DS_DIV_MAX, DS_VMAX, DS_UMAX, DS_VMAXC, DS_WMAX, DS_NAN_FLAG,
    DS_COUNT_FIXED = 32
};
// volume[l] at DS_COUNT_FIXED + l, init_volume[l] at DS_COUNT_FIXED + lmax + l

// Deterministic two-pass reductions: blocks write partials, one block folds them in order.
constexpr int RED_BLOCK = 256;
constexpr int RED_MAX_BLOCKS = 2048;

struct Profiler;

struct Ctx {
    icls_params P{};
    Grid g{};
    int lmax = 1;
    long nfield = 0;       // elements per device 3D field
    double NB2 = 0, rho0 = 0, visc = 0, pi = 0, dt = 0, time = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    bool initialised = false;
    bool first_pressure_update_done = false;

    // FP64 fields (C-halo/interior/LS all in the padded layout)
    double *unew = nullptr, *vnew = nullptr, *wnew = nullptr;
    double *dudt = nullptr, *dvdt = nullptr, *dwdt = nullptr;
    double *pnew = nullptr, *pold = nullptr, *p = nullptr, *p_jump = nullptr;
    double *duold = nullptr, *dvold = nullptr, *dwold = nullptr;
    double *p_x = nullptr, *p_y = nullptr, *p_z = nullptr;
    double *p_xold = nullptr, *p_yold = nullptr, *p_zold = nullptr;
    double *miu = nullptr, *rho = nullptr, *curv = nullptr, *curv2 = nullptr;
    double *ucc = nullptr, *vcc = nullptr, *wcc = nullptr;
    double *u_l = nullptr, *v_l = nullptr, *w_l = nullptr, *dheav = nullptr;
    double *ucc_l = nullptr, *vcc_l = nullptr, *wcc_l = nullptr;   // inflation velocities
    double *normal_m = nullptr;
    std::vector<double*> lvset;           // lmax level sets
    double *ls_a = nullptr, *ls_b = nullptr, *ls_c = nullptr;  // RK / re-init stage buffers
    uint8_t* mask = nullptr;              // narrow band (interior)
    double* thomas_d = nullptr;           // Thomas scratch (interior)
    double* dscal = nullptr;              // device scalars
    double* dpart = nullptr;              // reduction partials
    long npart_cap = 0;                   // capacity of dpart in doubles
    double** dls = nullptr;               // device table of the lmax level-set pointers
    double* stash = nullptr;              // two z planes (re-init residual reference)
    double* scratch = nullptr;            // one field for on-demand downloads
    int nscal = 0;                        // doubles in dscal / hscal
    double* hvol = nullptr;               // pinned copy of volume(l), init_volume(l)
    hipEvent_t ev_vol = nullptr;          // recorded after that copy
    bool vol_published = false;
    double* hscal = nullptr;              // pinned host mirror of dscal
    // constant tables
    double* pinvA = nullptr;              // 10 x 27 column-major
    double* xyrt = nullptr;               // nx * ny eigenvalues
    double* tw_x = nullptr;               // FFT twiddles along x (nx complex)
    double* tw_y = nullptr;               // along y
    std::vector<double> host_pinvA;

    Profiler* prof = nullptr;
};

// ---- error handling
void set_error(const char* fmt, ...);
#define HIP_OK(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::icls::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            return ICLS_ERR_HIP;                                                        \
        }                                                                               \
    } while (0)

// ---- profiling of named launches (HIP events on the library stream)
void prof_begin(Ctx& c, const char* name);
void prof_end(Ctx& c, const char* name);

struct ProfScope {
    Ctx& c;
    const char* n;
    ProfScope(Ctx& c_, const char* n_) : c(c_), n(n_) { if (c.prof) prof_begin(c, n); }
    ~ProfScope() { if (c.prof) prof_end(c, n); }
};

inline dim3 grid3(const Grid& g, int kcount, dim3 blk) {
    return dim3((g.nx + blk.x - 1) / blk.x, (g.ny + blk.y - 1) / blk.y, kcount);
}

// ---- launchers (defined in the k_*.hip files); each returns ICLS_OK or an error
// level set
int launch_tube(Ctx& c, const double* phi_in, double* phi_out);
int launch_bc_ls(Ctx& c, double* phi);
int launch_adv_stage(Ctx& c, int scheme, int stage, const double* phi0, const double* phis, double* out,
                     const double* vu, const double* vv, const double* vw, int vel_mode, const double* kappa,
                     const double* flag);
int launch_reinit_stage(Ctx& c, int stage, const double* phi0, const double* phis, double* out);
int launch_reinit_residual(Ctx& c, const double* phi0, const double* phinew);
int launch_boundmls(Ctx& c, double* phi);
int launch_init_levelset(Ctx& c, int l, double* phi);
// materials / reductions
int launch_materials(Ctx& c);
int launch_sum_interior(Ctx& c, const double* f, int slot);
int launch_inflation_prepare(Ctx& c, int l, int istep);
int launch_inflation_velocity(Ctx& c, int l);
// interface / gfm
int launch_curvature(Ctx& c, const double* phi, double* kappa, double* normal_m, const double* flag);
int launch_recurvature(Ctx& c, const double* phi, const double* kappa, double* kappa2);
int launch_gfm(Ctx& c, int l, const double* kappa2, bool fuse_fillps, bool store_pjump);
int launch_fillps(Ctx& c);
int launch_add_pjump(Ctx& c);
// poisson
int launch_fft_x(Ctx& c, double* f, int inverse);
int launch_fft_y(Ctx& c, double* f, int inverse);
int launch_thomas(Ctx& c, double* f);
// momentum / correction
int launch_momentum(Ctx& c);
int launch_bc_uvw(Ctx& c, double* u, double* v, double* w);
int launch_bc_p(Ctx& c, double* p);
int launch_correc(Ctx& c);
int launch_pressure_norm(Ctx& c);
int launch_pressure_rotate(Ctx& c, int first_step);
int launch_avg_vel(Ctx& c);
int launch_nan_check(Ctx& c);
int launch_chkdiv(Ctx& c);
int launch_chkdt_umax(Ctx& c);
int launch_max_vel(Ctx& c);
int launch_init_velocity(Ctx& c);
int launch_fill(Ctx& c, double* f, double v);
int launch_copy_interior(Ctx& c, const double* a, double* b);
int launch_momentum_rhs_only(Ctx& c);
int launch_jiemian(Ctx& c, double* out);
int launch_bc_p_pred(Ctx& c, double* p, const double* flag);
int launch_bc_uvw_pred(Ctx& c, double* u, double* v, double* w, const double* flag);
int launch_bc_ls_pred(Ctx& c, double* phi, const double* flag);
int launch_scalar_kernel(Ctx& c, int op, int a, int b);

}  // namespace icls
