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
//
// Precision classes.  "Field" arrays (level sets, velocities, momentum history,
// materials, curvature, GFM jumps) hold the mode's element type T: double in the FP64
// mode, float in the mixed modes.  "Pressure" arrays (pnew, pold, p, the Thomas
// multipliers) are double in every mode.  Context members are raw device addresses
// typed double*; launchers reinterpret field arrays as T*.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <cmath>
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
    DS_RES_MAX, DS_PN_A, DS_PN_B, DS_COUNT_FIXED = 32
};
// volume[l] at DS_COUNT_FIXED + l, init_volume[l] at DS_COUNT_FIXED + lmax + l

struct Profiler;

// Learned curvature closure: a 30 -> 64 -> 64 -> 1 tanh MLP on the 3 x 3 x 3 phi
// stencil (over dz) and the unit normal, giving kappa * dz.  Parameters are one flat
// FP32 vector: W1 (64 x 32, input padded to 32), b1, W2 (64 x 64), b2, w3, b3.
constexpr int MLP_NIN = 30, MLP_NPAD = 32, MLP_NH = 64;
constexpr int MLP_OW1 = 0;
constexpr int MLP_OB1 = MLP_OW1 + MLP_NH * MLP_NPAD;
constexpr int MLP_OW2 = MLP_OB1 + MLP_NH;
constexpr int MLP_OB2 = MLP_OW2 + MLP_NH * MLP_NH;
constexpr int MLP_OW3 = MLP_OB2 + MLP_NH;
constexpr int MLP_OB3 = MLP_OW3 + MLP_NH;
constexpr int MLP_NPARAM = MLP_OB3 + 1;
constexpr int MLP_GBLOCKS = 64;                 // blocks of the gradient kernel

struct Mlp {
    float* w = nullptr;                 // master weights
    float *adam_m = nullptr, *adam_v = nullptr;
    uint16_t* w1b = nullptr;            // bf16 W1 (64 x 32), W2 (64 x 64) and W2^T
    uint16_t* w2b = nullptr;
    uint16_t* w2tb = nullptr;
    float* gpart = nullptr;             // MLP_GBLOCKS x (MLP_NPARAM + 1) gradient partials
    float* grad = nullptr;              // reduced gradient (MLP_NPARAM) and loss
    double* pinvA_fine = nullptr;       // least-squares matrix at the 4x finer spacing
    int* list = nullptr;                // under-resolved interface cells of one level set
    int* list_t = nullptr;              // batched windows: the window of each listed cell
    int* count = nullptr;
    int list_cap = 0;
    int* total = nullptr;               // cells listed over all level sets in the last gfm pass
    long adam_t = 0;
    bool ready = false;                 // weights set or trained
    float gate = 0.25f;                 // |kappa| dz at which the closure is consulted
    float tol = 0.5f;                   // accepted |kappa_mlp - kappa_fd| dz
};

// The part of the grid one level set is stepped on: cells ox .. ox+wx-1 in x (and
// likewise y), layers oz+1 .. oz+wz in z.  wx is a power of two, wy a multiple of 8,
// and either the window does not cross the periodic boundary or it spans the whole
// grid.
struct LsWindow {
    int ox = 0, oy = 0, oz = 0, wx = 0, wy = 0, wz = 0;
};

// Several level-set windows in one launch (the drops of one colour: their windows do
// not overlap).  Window t owns blocks first[t] .. first[t+1]-1; a pointer argument
// given as nullptr stands for window t's own level set ls[t].  Every pointer is moved
// to the window's first cell, as a one-window launch sees it.
constexpr int WB_MAX = 64;
struct WinBatch {
    int n = 0;
    int first[WB_MAX + 1] = {};
    int ox[WB_MAX] = {}, oy[WB_MAX] = {}, oz[WB_MAX] = {}, wx[WB_MAX] = {}, wy[WB_MAX] = {}, wz[WB_MAX] = {};
    const void* ls[WB_MAX] = {};
    int lidx[WB_MAX] = {};      // level-set index of window t (its residual slot)
};
// the table travels in the kernel arguments, which hold 4 KiB with the other arguments
static_assert(sizeof(WinBatch) <= 3072, "WinBatch too large for the kernel argument segment");

struct Ctx {
    icls_params P{};
    Grid g{};
    Grid gfull{};                         // the whole grid (g is a window's inside a WindowScope)
    long ncell_all = 0;                   // interior cells of the whole grid
    std::vector<LsWindow> win;            // per level set; empty: every level set on the whole grid
    int win_age = 0;                      // ls_main passes since the windows were set
    int* dext = nullptr;                  // device extents, 6 ints per level set
    int* hext = nullptr;                  // pinned host copy
    int* dwin = nullptr;                  // device copy of the windows, 6 ints per level set
    int* hwin = nullptr;                  // pinned staging of it
    // Windows of one colour do not overlap (ghost layers included): their level sets run
    // in one launch per kernel (WinBatch).
    std::vector<int> colour;              // per level set
    int ncolour = 0;
    int res_slot = 0;                     // dscal slot of the re-init residual (DS_REINIT_RES)
    int lmax = 1;
    long nfield = 0;       // elements per device 3D field
    bool fp32 = false;     // field arrays hold float (mixed modes)
    size_t esz = 8;        // bytes per field element
    double NB2 = 0, rho0 = 0, visc = 0, pi = 0, dt = 0, time = 0;
    double dt_old = 0;                    // dt of the previous step (variable-step AB2)
    hipStream_t stream = nullptr;
    int device = 0;
    int ncu = 0;                          // compute units of the device
    std::map<const void*, int> zchunk;    // z-chunk length per z-marching kernel
    std::map<const void*, int> resident;  // resident blocks per CU per looping kernel
    bool initialised = false;

    // field class (T)
    double *unew = nullptr, *vnew = nullptr, *wnew = nullptr;
    double *dudt = nullptr, *dvdt = nullptr, *dwdt = nullptr;
    double *p_jump = nullptr;
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
    double* frhs = nullptr;               // Poisson right-hand side (the array p itself in FP64)
    double* pf = nullptr;                 // mixed: float work array of the transforms
    double* pf0 = nullptr;                // mixed: float starting solution of the refinement pass
    bool pf0_step = false;                // pf0 holds the time step's last solution rounded to float
    // pressure class (double)
    double *pnew = nullptr, *pold = nullptr, *p = nullptr;
    // Lazy pressure rotation (main.f90:164-183).  After a step the pnew buffer holds
    // the raw solution p^n and the pold buffer the raw buffer of the step before; the
    // values the reference stores are formed from them where they are read (the
    // momentum kernel, downloads) with the reference's own operations:
    //   pnew = A - S_A,  po = (p_prev_lazy ? B - S_B : B),
    //   pold = (p_first ? po - S_A : po + S_B - S_A),  S_A = DS_PN_A, S_B = DS_PN_B.
    bool p_lazy = false;
    bool p_prev_lazy = false;
    bool p_first = false;
    bool p_in_pnew = false;               // the reference's p is the pnew buffer
    bool vbulk_fresh = false;             // DS_V_BULK is the interior sum of the current v
    // Face-jump masks: bit 0/1/2 of jmask says p_x/p_y/p_z at the cell may be nonzero
    // (jmask_old likewise for the old jumps).  Readers take a jump value only where its
    // bit is set; elsewhere the jump is zero whatever the array holds.
    uint8_t* jmask = nullptr;
    uint8_t* jmask_old = nullptr;
    double* zplane = nullptr;             // nx * ny zeros, never written: the masked-off jump reads
    double* thomas_d = nullptr;           // serial Thomas: multipliers (interior)
    double* thomas_y = nullptr;           // serial Thomas, mixed: forward-sweep values
    int thomas_seg = 0;                   // partitioned Thomas: segment length (0: serial)
    double* thomas_seed = nullptr;        // partitioned Thomas: d below each segment
    double* scratch = nullptr;            // one double array for on-demand downloads
    uint8_t* mask = nullptr;              // narrow band (interior)
    double* dscal = nullptr;              // device scalars
    double* dpart = nullptr;              // reduction partials
    long npart_cap = 0;                   // capacity of dpart in doubles
    double** dls = nullptr;               // device table of the lmax level-set pointers
    double* stash = nullptr;              // two z planes (re-init residual reference)
    int nscal = 0;                        // doubles in dscal / hscal
    double* hvol = nullptr;               // pinned copy of volume(l), init_volume(l)
    hipEvent_t ev_vol = nullptr;          // recorded after that copy
    bool vol_published = false;
    double* hscal = nullptr;              // pinned host mirror of dscal
    // constant tables (double, and float copies for the mixed modes)
    double* pinvA = nullptr;              // 10 x 27 column-major
    double* xyrt = nullptr;               // nx * ny eigenvalues
    double* tw_x = nullptr;               // FFT twiddles along x (nx/2 complex)
    double* tw_y = nullptr;               // along y
    double* stash_b = nullptr;            // batched windows: 2 stash planes per batch slot
    int res_b0 = 0;                       // batched windows: first of lmax residual slots
    int save0 = 0;                        // two scalar slots: maxchk and the inflation flag kept aside
    double* cos_tab = nullptr;            // solver2d_cos: z transform matrices, xrt, zrt (on first use)
    double* pinvA_f = nullptr;            // float copies (raw float arrays)
    double* tw_x_f = nullptr;
    double* tw_y_f = nullptr;
    std::vector<double> host_pinvA;

    Mlp mlp;
    bool learned = false;                 // ICLS_MODE_MIXED_LEARNED
    Profiler* prof = nullptr;
};

template <class T>
inline T* as(double* p) { return reinterpret_cast<T*>(p); }
template <class T>
inline const T* as(const double* p) { return reinterpret_cast<const T*>(p); }

// ---- error handling
void set_error(const char* fmt, ...);

// Layers per block for a z-marching kernel `fn` launched with `threads` threads per block
// over tiles_xy (x, y) tiles: a multiple of `mult` chosen so that the launch comes as
// close as possible to whole rounds of resident blocks (occupancy x CUs at once; a
// partly empty last round costs a full one), each block costing its layers plus `halo`
// prologue layers.  Cached per kernel.
int zchunk_for(Ctx& c, const void* fn, int threads, int tiles_xy, int mult, int halo);
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

// four consecutive elements (16-byte aligned) as one or two 16-byte vector accesses
template <class T>
__device__ __forceinline__ void ld4(const T* a, T* v) {
    if constexpr (sizeof(T) == 4) {
        const float4 t = *reinterpret_cast<const float4*>(a);
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else {
        const double2 t0 = *reinterpret_cast<const double2*>(a), t1 = *reinterpret_cast<const double2*>(a + 2);
        v[0] = t0.x; v[1] = t0.y; v[2] = t1.x; v[3] = t1.y;
    }
}
template <class T>
__device__ __forceinline__ void st4(T* a, const T* v) {
    if constexpr (sizeof(T) == 4) {
        *reinterpret_cast<float4*>(a) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
        *reinterpret_cast<double2*>(a) = make_double2(v[0], v[1]);
        *reinterpret_cast<double2*>(a + 2) = make_double2(v[2], v[3]);
    }
}

// XCD-aware block order.  The hardware hands consecutive workgroup ids round-robin to
// the 8 XCDs, each with its own L2.  Kernels that share halos between neighbouring
// blocks launch a 1D grid and take their tile from xcd_tile(): XCD x then works
// through one contiguous range of tile ids, so blocks that run side by side on an
// XCD are spatial neighbours and their halos meet in the same L2.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int per = n / NUM_XCD, rem = n % NUM_XCD;
    const int x = b % NUM_XCD, i = b / NUM_XCD;
    return x * per + (x < rem ? x : rem) + i;
}
// The same with the tile ids dealt to the XCDs in groups of G: XCD x takes groups x,
// x + 8, x + 16, ... so that every XCD holds a share of each region of the grid, which
// balances kernels whose work sits in a few regions (the narrow band).  A count that is
// not a multiple of 8 G keeps the contiguous ranges.
__device__ __forceinline__ int xcd_tile_dealt(int b, int n, int G) {
    if (G <= 0 || n % (NUM_XCD * G) != 0) return xcd_tile(b, n);
    const int x = b % NUM_XCD, i = b / NUM_XCD;
    return ((i / G) * NUM_XCD + x) * G + i % G;
}

// ---- launchers (defined in the k_*.hip files); each returns ICLS_OK or an error.
// Field-class pointers are reinterpreted as the mode's element type.
// level set
int noalloc_setup(int device);
int chkdiv_noalloc(const void* u, const void* v, const void* w, int elem_bytes, const Grid& g, double* out,
                   hipStream_t stream);
int launch_tube(Ctx& c, const double* phi_in, double* phi_out);
int launch_ls_extent(Ctx& c, const double* phi, int* ext);
// the launch grid is a level-set window, not the whole grid
// the ls_mass_correction trigger on |V/V0 - 1| (levelset.f90:392): 1e-5, or 10^-n with
// mass_trigger_exp = n > 0
inline double mass_trigger(const icls_params& P) {
    return P.mass_trigger_exp > 0 ? std::pow(10., -P.mass_trigger_exp) : 1e-5;
}
inline bool in_window(const Ctx& c) { return c.g.nx != c.gfull.nx || c.g.ny != c.gfull.ny || c.g.nz != c.gfull.nz; }
// the window of block b of a batched launch: its grid, the offset of its first cell and
// of its first (i, j) column, its index t and the block's index inside the window
struct WinPick {
    Grid g;
    long off, poff;
    int t, bid;
};
__device__ __forceinline__ WinPick win_pick(const Grid& g, const WinBatch& wb, int b) {
    int t = 0;
    while (t + 1 < wb.n && b >= wb.first[t + 1]) t++;
    WinPick w;
    w.g = g;
    w.g.nx = wb.wx[t]; w.g.ny = wb.wy[t]; w.g.nz = wb.wz[t];
    w.poff = (long)wb.ox[t] + g.sy * (long)wb.oy[t];
    w.off = w.poff + g.sz * (long)wb.oz[t];
    w.t = t;
    w.bid = b - wb.first[t];
    return w;
}
template <class P>
__device__ __forceinline__ P* win_ptr(P* shared, const WinBatch& wb, const WinPick& w) {
    return (shared ? shared : (P*)wb.ls[w.t]) + w.off;
}
// The level sets ls (windows of one colour, which do not overlap) in launches of at
// most WB_MAX windows, each launch holding windows with the same key(window); blocks
// (window) is the window's block count.  launch(batch, key, members) issues a launch.
template <class KeyFn, class BlocksFn, class LaunchFn>
inline int for_each_batch(Ctx& c, const std::vector<int>& ls, KeyFn key, BlocksFn blocks, LaunchFn launch) {
    std::vector<int> keys;
    for (int l : ls) {
        const int k = key(c.win[l]);
        bool seen = false;
        for (int q : keys) seen = seen || q == k;
        if (!seen) keys.push_back(k);
    }
    for (int k : keys) {
        WinBatch wb;
        std::vector<int> members;
        auto flush = [&]() -> int {
            if (wb.n == 0) return ICLS_OK;
            int rc = launch(wb, k, members);
            wb = WinBatch();
            members.clear();
            return rc;
        };
        for (int l : ls) {
            const LsWindow& w = c.win[l];
            if (key(w) != k) continue;
            const long nb = blocks(w);
            if (nb <= 0 || (long)wb.first[wb.n] + nb > (1L << 30)) {
                set_error("window batch: bad block count");
                return ICLS_ERR_STATE;
            }
            const int t = wb.n;
            wb.ox[t] = w.ox; wb.oy[t] = w.oy; wb.oz[t] = w.oz;
            wb.wx[t] = w.wx; wb.wy[t] = w.wy; wb.wz[t] = w.wz;
            wb.ls[t] = c.lvset[l];
            wb.lidx[t] = l;
            wb.first[t + 1] = wb.first[t] + (int)nb;
            wb.n = t + 1;
            members.push_back(l);
            if (wb.n == WB_MAX) { int rc = flush(); if (rc) return rc; }
        }
        int rc = flush();
        if (rc) return rc;
    }
    return ICLS_OK;
}

// one colour's windows (level sets ls), batched: tube + the three advection stages, and
// the three re-initialisation stages; the re-init residual of level set l goes to
// dscal[res_b0 + l], and that of level set lmax - 1 also to DS_REINIT_RES
int launch_advect_batch(Ctx& c, const std::vector<int>& ls);
int launch_reinit_batch(Ctx& c, const std::vector<int>& ls);
int launch_adv_stage(Ctx& c, int scheme, int stage, const double* phi0, const double* phis, double* out,
                     const double* vu, const double* vv, const double* vw, int velf, const double* flag);
int launch_reinit_stage(Ctx& c, int stage, const double* phi0, const double* phis, double* out);
int launch_boundmls(Ctx& c, double* phi);
int launch_init_levelset(Ctx& c, int l, double* phi);
// materials / reductions
int launch_materials(Ctx& c);
int launch_jiemian(Ctx& c, double* out);
int launch_sum_interior(Ctx& c, const double* f, int slot);
int launch_inflation_prepare(Ctx& c, int l, int istep);
int launch_inflation_velocity(Ctx& c, int l);
int launch_ipot(Ctx& c, double* dtemp);
// interface / gfm
int launch_curvature(Ctx& c, const double* phi, double* kappa, double* normal_m, const double* flag,
                     bool listed = false);
int launch_recurvature(Ctx& c, const double* phi, const double* kappa, double* kappa2);
int launch_gfm(Ctx& c, int l, const double* kappa2, bool fuse_fillps, bool store_pjump);
int launch_fillps(Ctx& c);
int launch_add_pjump(Ctx& c);
// poisson: solve L p = frhs into p (solver2d; in the mixed modes with one refinement,
// which with warm starts from the previous time step's solution where one is held)
int launch_solve(Ctx& c, bool warm);
// solver2d_cos: p -> solution with y rule 1 Walls, 2 InOut, 3 OutOut, in double in every mode
int launch_solve_cos(Ctx& c, int bcy);
int launch_pressure_materialize(Ctx& c);
// pressure norm, lazy rotation and NaN check of a time step's end (main.f90:164-185, 207)
int launch_step_end(Ctx& c, int first_step);
int launch_jump_densify(Ctx& c, bool old);   // zero the unmasked jumps, mask every cell
int launch_pressure_release_p(Ctx& c);   // the work array p takes the raw solution back
int thomas_seg_len(int nz);
int launch_thomas_seed(Ctx& c);
// learned curvature closure (k_mlp.hip)
int mlp_alloc(Ctx& c);
void mlp_free(Ctx& c);
int mlp_set_weights(Ctx& c, const float* w);
int mlp_forward(Ctx& c, const uint16_t* x_bf16, int n, float* y);          // device pointers
int mlp_grad(Ctx& c, const uint16_t* x_bf16, const float* t, int n, float* grad_loss);
int mlp_adam(Ctx& c, float lr);                                            // after mlp_grad
int mlp_samples(Ctx& c, long n, unsigned long long seed, float r_min, float r_max, float gate, uint16_t* x_bf16,
                float* t, float* fd);
int launch_mlp_reset_list(Ctx& c);
int launch_mlp_reset_list_total(Ctx& c);   // both counters in one launch
int launch_mlp_curv(Ctx& c, const double* phi, const double* kappa_fd, double* kappa_out);
// the closure over the cells listed by a batched curvature launch (whole-grid indices,
// each with its window in the batch wb)
int launch_mlp_curv_batch(Ctx& c, const WinBatch& wb, const double* kappa_fd, double* kappa_out);
// get_curvature, re_curvature, the closure (learned) and the added GFM jumps of the
// level sets ls, windows of one colour, in batched launches
int launch_curvature_jumps_batch(Ctx& c, const std::vector<int>& ls, bool learned);
// momentum / correction
int launch_momentum(Ctx& c);
int launch_momentum_rhs_only(Ctx& c);
int launch_bc_uvw(Ctx& c, double* u, double* v, double* w);
int launch_bc_uvw_pred(Ctx& c, double* u, double* v, double* w, const double* flag);
int launch_bc_p(Ctx& c, double* p);
int launch_bc_p_t_pred(Ctx& c, double* p, const double* flag);
int launch_correc(Ctx& c);
int launch_pressure_norm(Ctx& c);
int launch_pressure_rotate(Ctx& c, int first_step);
int launch_avg_vel(Ctx& c);
int launch_nan_check(Ctx& c);
int launch_chkdiv(Ctx& c);
int launch_chkdt_umax(Ctx& c);
int launch_max_vel(Ctx& c);
int launch_init_velocity(Ctx& c);
int launch_fill(Ctx& c, double* f, double v, int fp32);
int launch_copy_field(Ctx& c, const double* a, double* b);
int launch_convert(Ctx& c, const double* src, int src_fp32, double* dst, int dst_fp32);

}  // namespace icls
