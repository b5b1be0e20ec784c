// k_interface.hip -- least-squares curvature (get_curvature), its smoothing
// (re_curvature), the ghost-fluid pressure jumps (gfm4p) and the Poisson right-hand
// side (fillps), fused where the reference applies them back to back.
//
// T is the element type of the field arrays; the right-hand side the Poisson solve
// reads (c.frhs) has the same type.  In the FP64 mode c.frhs is the pressure array p
// itself, as in the reference.
#include <initializer_list>

#include "icls_internal.hpp"
#include "icls_stencil.hpp"

namespace icls {

static constexpr int BX = 64, BY = 4;

// 27-point gather in reshape order (i fastest, then j, then k), interface.f90:59
template <class T>
__device__ __forceinline__ void gather27(const Grid& g, const T* __restrict__ f, int i0, int j0, int k, T* v) {
    int im = wrap(i0 - 1, g.nx), ip = wrap(i0 + 1, g.nx), jm = wrap(j0 - 1, g.ny), jp = wrap(j0 + 1, g.ny);
    int is[3] = {im, i0, ip}, js[3] = {jm, j0, jp};
    int q = 0;
#pragma unroll
    for (int c = -1; c <= 1; c++)
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int a = 0; a < 3; a++) v[q++] = f[gidx(g, is[a], js[b], k + c)];
}

// row r of matmul(pseudo_inv_A, v), accumulated over the 27 columns in order
template <class T>
__device__ __forceinline__ T lsq_row(const T* __restrict__ A, int r, const T* v) {
    T s = T(0.);
#pragma unroll
    for (int q = 0; q < 27; q++) s = s + A[r + 10 * q] * v[q];
    return s;
}

// get_curvature (interface.f90:20-119): for |phi| <= NarrowBand_2 - dz fit the
// quadratic of init_matrix to the 27 values, kappa = -div(grad phi/|grad phi|) with
// the 0.9 gradient filter and the +-1/dz and 1e-9 limits; zero elsewhere.  The
// z-ghost layers of kappa are never written and stay zero (interface.f90:49).
// With a cell list (the learned mode) the cells of the re_curvature shell whose
// curvature is under-resolved -- |phi| < 2 dz, |grad phi| >= 0.9 and |kappa| dz >= gate
// -- are appended to it for the learned closure (k_mlp_curv).
// A listed cell is entered as gofs + i0 + sy j0 + sz (k - 1): gofs is 0 for a launch on
// one grid or window (the index is the window's own), and the window's offset in the
// whole grid for a batched launch, which also enters the window t in list_t.
// the curvature of a band cell (phi value ph) and its gradient filter m
template <class T>
__device__ __forceinline__ T curvature_band(const Grid& g, const T* __restrict__ A, const T* __restrict__ phi, T ph,
                                            T& m, int* __restrict__ list, int* __restrict__ list_t, int t, long gofs,
                                            int* __restrict__ count, int cap, T gate, int i0, int j0, int k) {
    const T dz = T(g.dz), max_kappa = T(1.) / dz;
    T v[27];
    gather27(g, phi, i0, j0, k, v);
    const T cv = curv_lsq(A, v, max_kappa, m);
    if (list && fabs(ph) < T(2.) * dz && m >= T(0.9) && fabs(cv) * dz >= gate) {
        const int slot = atomicAdd(count, 1);
        if (slot < cap) {
            list[slot] = (int)(gofs + i0 + g.sy * j0 + g.sz * (long)(k - 1));
            if (list_t) list_t[slot] = t;
        }
    }
    return cv;
}

template <class T>
__device__ __forceinline__ void curvature_at(const Grid& g, T NB2, const T* __restrict__ A, const T* __restrict__ phi,
                                             T* __restrict__ kappa, T* __restrict__ nm, int* __restrict__ list,
                                             int* __restrict__ list_t, int t, long gofs, int* __restrict__ count,
                                             int cap, T gate, int i0, int j0, int k) {
    long q = gidx(g, i0, j0, k);
    const T dz = T(g.dz), ph = phi[q];
    T cv = T(0.), m = T(1.);
    if (fabs(ph) <= NB2 - T(1.) * dz)
        cv = curvature_band<T>(g, A, phi, ph, m, list, list_t, t, gofs, count, cap, gate, i0, j0, k);
    kappa[q] = cv;
    if (nm) nm[q] = m;
}

template <class T>
__global__ void k_curvature(Grid g, T NB2, const T* __restrict__ A, const T* __restrict__ phi, T* __restrict__ kappa,
                            T* __restrict__ nm, const double* __restrict__ flag, int* __restrict__ list,
                            int* __restrict__ count, int cap, T gate) {
    if (flag && flag[0] == 0.) return;
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    curvature_at<T>(g, NB2, A, phi, kappa, nm, list, nullptr, 0, 0, count, cap, gate, i0, j0, k);
}

// the same with four consecutive cells of a row per thread (nx a multiple of 256):
// phi, kappa and the normal magnitude move as 16-byte vectors
template <class T>
__global__ void __launch_bounds__(BX * BY)
k_curvature4(Grid g, T NB2, const T* __restrict__ A, const T* __restrict__ phi, T* __restrict__ kappa,
             T* __restrict__ nm, const double* __restrict__ flag, int* __restrict__ list, int* __restrict__ count,
             int cap, T gate) {
    if (flag && flag[0] == 0.) return;
    const int i0 = 4 * (blockIdx.x * BX + threadIdx.x), j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (j0 >= g.ny) return;
    const long q = gidx(g, i0, j0, k);
    const T dz = T(g.dz);
    T ph[4], cv[4], m[4];
    ld4(phi + q, ph);
#pragma unroll
    for (int t = 0; t < 4; t++) {
        cv[t] = T(0.);
        m[t] = T(1.);
        if (fabs(ph[t]) <= NB2 - T(1.) * dz)
            cv[t] = curvature_band<T>(g, A, phi, ph[t], m[t], list, nullptr, 0, 0, count, cap, gate, i0 + t, j0, k);
    }
    st4(kappa + q, cv);
    if (nm) st4(nm + q, m);
}

// ---------------------------------------------------------------- batched windows
// A launch over the windows of a batch: 64 x 4 tiles, one layer each, window t owning
// blocks wb.first[t] .. wb.first[t+1]-1; the cell's window coordinates, and the window.
struct WinCell {
    WinPick w;
    int i0, j0, k;
    bool in;
};
__device__ __forceinline__ WinCell win_cell(const Grid& g, const WinBatch& wb) {
    WinCell r;
    r.w = win_pick(g, wb, blockIdx.x);
    const int ntx = (r.w.g.nx + BX - 1) / BX, nty = (r.w.g.ny + BY - 1) / BY;
    const int b = r.w.bid;
    r.i0 = (b % ntx) * BX + threadIdx.x;
    r.j0 = ((b / ntx) % nty) * BY + threadIdx.y;
    r.k = b / (ntx * nty) + 1;
    r.in = r.i0 < r.w.g.nx && r.j0 < r.w.g.ny;
    return r;
}
static long win_blocks(const LsWindow& w) { return (long)((w.wx + BX - 1) / BX) * ((w.wy + BY - 1) / BY) * w.wz; }

template <class T>
__global__ void k_curvature_batch(Grid g, WinBatch wb, T NB2, const T* __restrict__ A, T* __restrict__ kappa,
                                  int* __restrict__ list, int* __restrict__ list_t, int* __restrict__ count, int cap,
                                  T gate) {
    const WinCell wc = win_cell(g, wb);
    if (!wc.in) return;
    const int t = wc.w.t;
    const long gofs = (long)wb.ox[t] + g.sy * (long)wb.oy[t] + g.sz * (long)wb.oz[t];
    curvature_at<T>(wc.w.g, NB2, A, (const T*)wb.ls[t] + wc.w.off, kappa + wc.w.off, (T*)nullptr, list, list_t, t,
                    gofs, count, cap, gate, wc.i0, wc.j0, wc.k);
}

// the four-cells-per-thread kernels: whole-grid launches (inside a level-set window
// the band fills most cells, and one cell per thread measured faster there), rows of a
// multiple of 256 cells, and the vector arrays 16-byte aligned at every row
static bool vec4_ok(const Ctx& c, std::initializer_list<const void*> ptrs) {
    const Grid& g = c.g;
    if (in_window(c) || g.nx % (4 * BX) != 0 || g.sy % 4 != 0 || g.sz % 4 != 0) return false;
    for (const void* p : ptrs)
        if (p && reinterpret_cast<uintptr_t>(p) % 16 != 0) return false;
    return true;
}

int launch_curvature(Ctx& c, const double* phi, double* kappa, double* normal_m, const double* flag, bool listed) {
    ProfScope ps(c, "k_curvature");
    dim3 blk(BX, BY, 1);
    dim3 grd = grid3(c.g, c.g.nz, blk);
    int* list = listed ? c.mlp.list : nullptr;
    if (vec4_ok(c, {phi, kappa, normal_m})) {
        grd.x = c.g.nx / (4 * BX);
        if (c.fp32)
            hipLaunchKernelGGL(k_curvature4<float>, grd, blk, 0, c.stream, c.g, (float)c.NB2, as<float>(c.pinvA_f),
                               as<float>(phi), as<float>(kappa), as<float>(normal_m), flag, list, c.mlp.count,
                               c.mlp.list_cap, c.mlp.gate);
        else
            hipLaunchKernelGGL(k_curvature4<double>, grd, blk, 0, c.stream, c.g, c.NB2, (const double*)c.pinvA, phi,
                               kappa, normal_m, flag, list, c.mlp.count, c.mlp.list_cap, (double)c.mlp.gate);
    } else if (c.fp32)
        hipLaunchKernelGGL(k_curvature<float>, grd, blk, 0, c.stream, c.g, (float)c.NB2, as<float>(c.pinvA_f),
                           as<float>(phi), as<float>(kappa), as<float>(normal_m), flag, list, c.mlp.count,
                           c.mlp.list_cap, c.mlp.gate);
    else
        hipLaunchKernelGGL(k_curvature<double>, grd, blk, 0, c.stream, c.g, c.NB2, (const double*)c.pinvA, phi, kappa,
                           normal_m, flag, list, c.mlp.count, c.mlp.list_cap, (double)c.mlp.gate);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// re_curvature (interface.f90:186-218): for |phi| < 2 dz the constant term of the
// least-squares fit to the 27 curvature values; zero elsewhere.  This smoothed value is
// what the GFM jumps interpolate to the interface.
template <class T>
__device__ __forceinline__ void recurvature_at(const Grid& g, const T* __restrict__ A, const T* __restrict__ phi,
                                               const T* __restrict__ kappa, T* __restrict__ kappa2, int i0, int j0,
                                               int k) {
    long q = gidx(g, i0, j0, k);
    T r = T(0.);
    if (fabs(phi[q]) < T(2.) * T(g.dz)) {
        T v[27];
        gather27(g, kappa, i0, j0, k, v);
        r = lsq_row(A, 0, v);
    }
    kappa2[q] = r;
}

template <class T>
__global__ void k_recurvature_batch(Grid g, WinBatch wb, const T* __restrict__ A, const T* __restrict__ kappa,
                                    T* __restrict__ kappa2) {
    const WinCell wc = win_cell(g, wb);
    if (!wc.in) return;
    recurvature_at<T>(wc.w.g, A, (const T*)wb.ls[wc.w.t] + wc.w.off, kappa + wc.w.off, kappa2 + wc.w.off, wc.i0,
                      wc.j0, wc.k);
}

template <class T>
__global__ void k_recurvature(Grid g, const T* __restrict__ A, const T* __restrict__ phi, const T* __restrict__ kappa,
                              T* __restrict__ kappa2) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    recurvature_at<T>(g, A, phi, kappa, kappa2, i0, j0, k);
}

// the same with four consecutive cells of a row per thread (nx a multiple of 256)
template <class T>
__global__ void __launch_bounds__(BX * BY)
k_recurvature4(Grid g, const T* __restrict__ A, const T* __restrict__ phi, const T* __restrict__ kappa,
               T* __restrict__ kappa2) {
    const int i0 = 4 * (blockIdx.x * BX + threadIdx.x), j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (j0 >= g.ny) return;
    const long q = gidx(g, i0, j0, k);
    T ph[4], r[4];
    ld4(phi + q, ph);
#pragma unroll
    for (int t = 0; t < 4; t++) {
        r[t] = T(0.);
        if (fabs(ph[t]) < T(2.) * T(g.dz)) {
            T v[27];
            gather27(g, kappa, i0 + t, j0, k, v);
            r[t] = lsq_row(A, 0, v);
        }
    }
    st4(kappa2 + q, r);
}

int launch_recurvature(Ctx& c, const double* phi, const double* kappa, double* kappa2) {
    ProfScope ps(c, "k_recurvature");
    dim3 blk(BX, BY, 1);
    dim3 grd = grid3(c.g, c.g.nz, blk);
    if (c.fp32 && vec4_ok(c, {phi, kappa2})) {
        // float only: in FP64 the four cells' registers cost more occupancy than the
        // vector accesses gain
        grd.x = c.g.nx / (4 * BX);
        hipLaunchKernelGGL(k_recurvature4<float>, grd, blk, 0, c.stream, c.g, as<float>(c.pinvA_f), as<float>(phi),
                           as<float>(kappa), as<float>(kappa2));
    } else if (c.fp32)
        hipLaunchKernelGGL(k_recurvature<float>, grd, blk, 0, c.stream, c.g, as<float>(c.pinvA_f), as<float>(phi),
                           as<float>(kappa), as<float>(kappa2));
    else
        hipLaunchKernelGGL(k_recurvature<double>, grd, blk, 0, c.stream, c.g, (const double*)c.pinvA, phi, kappa, kappa2);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// fillps (fillps.f90:36-53): rho0/dt * div(u*)
template <class T>
__device__ __forceinline__ T fillps_rhs(const Grid& g, T coef, const T* __restrict__ du, const T* __restrict__ dv,
                                        const T* __restrict__ dw, int i0, int j0, int k) {
    long q = gidx(g, i0, j0, k);
    T vel_x = (du[q] - du[gidx(g, wrap(i0 - 1, g.nx), j0, k)]) * T(g.dxi);
    T vel_y = (dv[q] - dv[gidx(g, i0, wrap(j0 - 1, g.ny), k)]) * T(g.dyi);
    T vel_z = (dw[q] - dw[gidx(g, i0, j0, k - 1)]) * T(g.dzi);
    return coef * (vel_x + vel_y + vel_z);
}

// gfm4p jumps of one level set at one cell (gfm.f90:99-215).  Returns the separate
// jump Laplacian pj_sep and the + face jumps px/py/pz (already sign-flipped inside
// the drop); false when the cell is outside the |phi| <= 2 dz shell.
template <class T>
__device__ __forceinline__ bool gfm_cell(const Grid& g, T nondi, const T* __restrict__ ls, const T* __restrict__ cvf,
                                         int i0, int j0, int k, T& pjs, T& pxs, T& pys, T& pzs) {
    long q = gidx(g, i0, j0, k);
    T phc = ls[q];
    if (!(fabs(phc) <= T(2.) * T(g.dz))) return false;
    const T dx2 = T(g.dx) * T(g.dx), dy2 = T(g.dy) * T(g.dy), dz2 = T(g.dz) * T(g.dz);
    T ph[3], cv[3], jump[3];
    T face[3] = {T(0.), T(0.), T(0.)};
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        long qm, qp;
        if (ax == 0) { qm = gidx(g, wrap(i0 - 1, g.nx), j0, k); qp = gidx(g, wrap(i0 + 1, g.nx), j0, k); }
        else if (ax == 1) { qm = gidx(g, i0, wrap(j0 - 1, g.ny), k); qp = gidx(g, i0, wrap(j0 + 1, g.ny), k); }
        else { qm = gidx(g, i0, j0, k - 1); qp = gidx(g, i0, j0, k + 1); }
        ph[0] = ls[qm]; ph[1] = phc; ph[2] = ls[qp];
        cv[0] = cvf[qm]; cv[1] = cvf[q]; cv[2] = cvf[qp];
        T jm = T(0.), jp = T(0.);
        if (ph[1] * ph[0] < T(0.)) jm = nondi * (-icurvature<T>(-1, ph + 1, cv + 1));
        if (ph[1] * ph[2] < T(0.)) { jp = nondi * (-icurvature<T>(1, ph + 1, cv + 1)); face[ax] = jp; }
        jump[ax] = jm + T(0.) + jp;   // sum(jp_x) over (-1, 0, 1)
    }
    if (phc > T(0.)) {
        pjs = (jump[0] / dx2 + jump[1] / dy2 + jump[2] / dz2);
        pxs = face[0]; pys = face[1]; pzs = face[2];
    } else if (phc < T(0.)) {
        pjs = -(jump[0] / dx2 + jump[1] / dy2 + jump[2] / dz2);
        pxs = -face[0]; pys = -face[1]; pzs = -face[2];
    } else {
        pjs = T(0.); pxs = T(0.); pys = T(0.); pzs = T(0.);
    }
    return true;
}

// MODE 0: first level set -- write p_jump, p_x, p_y, p_z over the interior (zero off
//         the shell), gfm.f90:71-79 and 219-224.
// MODE 1: further level sets -- add inside their shell (gfm.f90:226-232).
// MODE 2: single level set fused with fillps and main.f90:142 -- rhs = fillps + p_jump.
template <class T, int MODE>
__global__ void k_gfm(Grid g, T nondi, T coef, const T* __restrict__ ls, const T* __restrict__ cvf,
                      T* __restrict__ pjump, T* __restrict__ px, T* __restrict__ py, T* __restrict__ pz,
                      const T* __restrict__ du, const T* __restrict__ dv, const T* __restrict__ dw,
                      T* __restrict__ rhs, uint8_t* __restrict__ jm) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    long q = gidx(g, i0, j0, k);
    T pjs = T(0.), pxs = T(0.), pys = T(0.), pzs = T(0.);
    bool shell = gfm_cell<T>(g, nondi, ls, cvf, i0, j0, k, pjs, pxs, pys, pzs);
    if (MODE == 1) {
        if (shell) {
            pjump[q] = pjump[q] + pjs;
            px[q] = px[q] + pxs; py[q] = py[q] + pys; pz[q] = pz[q] + pzs;
        }
        return;
    }
    if (MODE == 2) {
        // the face jumps are zero outside a few cells of the interface: only the nonzero
        // ones are stored, and the mask says which
        const unsigned bits = (pxs != T(0.) ? 1u : 0u) | (pys != T(0.) ? 2u : 0u) | (pzs != T(0.) ? 4u : 0u);
        jm[q] = (uint8_t)bits;
        if (bits & 1u) px[q] = pxs;
        if (bits & 2u) py[q] = pys;
        if (bits & 4u) pz[q] = pzs;
    } else {
        px[q] = pxs; py[q] = pys; pz[q] = pzs;
        jm[q] = 7;
    }
    if (pjump) pjump[q] = pjs;
    if (MODE == 2) rhs[q] = fillps_rhs<T>(g, coef, du, dv, dw, i0, j0, k) + pjs;
}

// MODE 2 with four consecutive cells of a row per thread on the whole grid (nx a
// multiple of 256): the mask, p_jump, the momentum results and the right-hand side move
// as vectors; per cell the operations of k_gfm and fillps_rhs
template <class T>
__global__ void __launch_bounds__(BX * BY)
k_gfm4(Grid g, T nondi, T coef, const T* __restrict__ ls, const T* __restrict__ cvf, T* __restrict__ pjump,
       T* __restrict__ px, T* __restrict__ py, T* __restrict__ pz, const T* __restrict__ du,
       const T* __restrict__ dv, const T* __restrict__ dw, T* __restrict__ rhs, uint8_t* __restrict__ jm) {
    const int i0 = 4 * (blockIdx.x * BX + threadIdx.x), j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (j0 >= g.ny) return;
    const long q = gidx(g, i0, j0, k);
    T pj[4];
    unsigned bits[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        T pjs = T(0.), pxs = T(0.), pys = T(0.), pzs = T(0.);
        gfm_cell<T>(g, nondi, ls, cvf, i0 + t, j0, k, pjs, pxs, pys, pzs);
        bits[t] = (pxs != T(0.) ? 1u : 0u) | (pys != T(0.) ? 2u : 0u) | (pzs != T(0.) ? 4u : 0u);
        if (bits[t] & 1u) px[q + t] = pxs;
        if (bits[t] & 2u) py[q + t] = pys;
        if (bits[t] & 4u) pz[q + t] = pzs;
        pj[t] = pjs;
    }
    *reinterpret_cast<uchar4*>(jm + q) =
        make_uchar4((uint8_t)bits[0], (uint8_t)bits[1], (uint8_t)bits[2], (uint8_t)bits[3]);
    if (pjump) st4(pjump + q, pj);
    T duc[4], dvc[4], dvm[4], dwc[4], dwm[4], r[4];
    ld4(du + q, duc);
    ld4(dv + q, dvc);
    ld4(dv + gidx(g, i0, wrap(j0 - 1, g.ny), k), dvm);
    ld4(dw + q, dwc);
    ld4(dw + q - g.sz, dwm);
    const T dum = du[gidx(g, wrap(i0 - 1, g.nx), j0, k)];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const T vel_x = (duc[t] - (t ? duc[t - 1] : dum)) * T(g.dxi);
        const T vel_y = (dvc[t] - dvm[t]) * T(g.dyi);
        const T vel_z = (dwc[t] - dwm[t]) * T(g.dzi);
        r[t] = coef * (vel_x + vel_y + vel_z) + pj[t];
    }
    st4(rhs + q, r);
}

// MODE 1 over the windows of a batch: each level set's jumps added inside its shell
template <class T>
__global__ void k_gfm_batch(Grid g, WinBatch wb, T nondi, const T* __restrict__ cvf, T* __restrict__ pjump,
                            T* __restrict__ px, T* __restrict__ py, T* __restrict__ pz) {
    const WinCell wc = win_cell(g, wb);
    if (!wc.in) return;
    const long o = wc.w.off;
    const long q = gidx(wc.w.g, wc.i0, wc.j0, wc.k) + o;
    T pjs = T(0.), pxs = T(0.), pys = T(0.), pzs = T(0.);
    if (gfm_cell<T>(wc.w.g, nondi, (const T*)wb.ls[wc.w.t] + o, cvf + o, wc.i0, wc.j0, wc.k, pjs, pxs, pys, pzs)) {
        pjump[q] = pjump[q] + pjs;
        px[q] = px[q] + pxs; py[q] = py[q] + pys; pz[q] = pz[q] + pzs;
    }
}

template <class T>
static void gfm_dispatch(Ctx& c, int l, const double* kappa2, int mode, bool store_pjump) {
    dim3 blk(BX, BY, 1);
    dim3 grd = grid3(c.g, c.g.nz, blk);
    const T nondi = (T)(1. / c.P.We), coef = (T)(c.rho0 / c.dt);
    const T* ls = as<T>(c.lvset[l]);
    const T* k2 = as<T>(kappa2);
    T *px = as<T>(c.p_x), *py = as<T>(c.p_y), *pz = as<T>(c.p_z);
    T* pj = store_pjump ? as<T>(c.p_jump) : (T*)nullptr;
    // float only: in FP64 the four cells' registers cost more than the vectors gain
    if (mode == 2 && c.fp32 && vec4_ok(c, {ls, pj, c.dudt, c.dvdt, c.dwdt, c.frhs})) {
        grd.x = c.g.nx / (4 * BX);
        hipLaunchKernelGGL(k_gfm4<T>, grd, blk, 0, c.stream, c.g, nondi, coef, ls, k2, pj, px, py, pz, as<T>(c.dudt),
                           as<T>(c.dvdt), as<T>(c.dwdt), as<T>(c.frhs), c.jmask);
    } else if (mode == 2)
        hipLaunchKernelGGL((k_gfm<T, 2>), grd, blk, 0, c.stream, c.g, nondi, coef, ls, k2,
                           store_pjump ? as<T>(c.p_jump) : (T*)nullptr, px, py, pz, as<T>(c.dudt), as<T>(c.dvdt),
                           as<T>(c.dwdt), as<T>(c.frhs), c.jmask);
    else if (mode == 0)
        hipLaunchKernelGGL((k_gfm<T, 0>), grd, blk, 0, c.stream, c.g, nondi, coef, ls, k2, as<T>(c.p_jump), px, py, pz,
                           (const T*)nullptr, (const T*)nullptr, (const T*)nullptr, (T*)nullptr, c.jmask);
    else
        hipLaunchKernelGGL((k_gfm<T, 1>), grd, blk, 0, c.stream, c.g, nondi, coef, ls, k2, as<T>(c.p_jump), px, py, pz,
                           (const T*)nullptr, (const T*)nullptr, (const T*)nullptr, (T*)nullptr, c.jmask);
}

// fuse_fillps: the single-level-set step (MODE 2); otherwise MODE 0 for l = 0 and
// MODE 1 for the further level sets
int launch_gfm(Ctx& c, int l, const double* kappa2, bool fuse_fillps, bool store_pjump) {
    ProfScope ps(c, "k_gfm");
    const int mode = fuse_fillps ? 2 : (l == 0 ? 0 : 1);
    if (c.fp32) gfm_dispatch<float>(c, l, kappa2, mode, store_pjump);
    else gfm_dispatch<double>(c, l, kappa2, mode, store_pjump);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// get_curvature, re_curvature, the learned closure (learned) and the added jumps of the
// level sets ls -- windows of one colour, which do not overlap -- one launch per kernel
// for up to WB_MAX windows.  Per cell the operations are those of the launches on one
// window at a time; the closure's cell list holds whole-grid indices and their windows.
template <class T>
static int curvature_jumps_batch_t(Ctx& c, const std::vector<int>& ls, bool learned) {
    auto key = [](const LsWindow&) { return 0; };
    auto blocks = [](const LsWindow& w) { return win_blocks(w); };
    const T nondi = (T)(1. / c.P.We);
    const T* A = as<T>(c.fp32 ? c.pinvA_f : c.pinvA);
    return for_each_batch(c, ls, key, blocks, [&](const WinBatch& wb, int, const std::vector<int>&) -> int {
        const dim3 grd((unsigned)wb.first[wb.n]), blk(BX, BY, 1);
        if (learned) {
            int rc = launch_mlp_reset_list(c);
            if (rc) return rc;
        }
        {
            ProfScope ps(c, "k_curvature");
            hipLaunchKernelGGL(k_curvature_batch<T>, grd, blk, 0, c.stream, c.g, wb, (T)c.NB2, A, as<T>(c.curv),
                               learned ? c.mlp.list : nullptr, learned ? c.mlp.list_t : nullptr, c.mlp.count,
                               c.mlp.list_cap, (T)c.mlp.gate);
            HIP_OK(hipGetLastError());
        }
        {
            ProfScope ps(c, "k_recurvature");
            hipLaunchKernelGGL(k_recurvature_batch<T>, grd, blk, 0, c.stream, c.g, wb, A, (const T*)as<T>(c.curv),
                               as<T>(c.curv2));
            HIP_OK(hipGetLastError());
        }
        if (learned) {
            int rc = launch_mlp_curv_batch(c, wb, c.curv, c.curv2);
            if (rc) return rc;
        }
        ProfScope ps(c, "k_gfm");
        hipLaunchKernelGGL(k_gfm_batch<T>, grd, blk, 0, c.stream, c.g, wb, nondi, (const T*)as<T>(c.curv2),
                           as<T>(c.p_jump), as<T>(c.p_x), as<T>(c.p_y), as<T>(c.p_z));
        HIP_OK(hipGetLastError());
        return ICLS_OK;
    });
}

int launch_curvature_jumps_batch(Ctx& c, const std::vector<int>& ls, bool learned) {
    return c.fp32 ? curvature_jumps_batch_t<float>(c, ls, learned) : curvature_jumps_batch_t<double>(c, ls, learned);
}

// every stored cell: jumps whose mask bit is clear become zero, every bit is set
template <class T>
__global__ void k_jump_densify(Grid g, uint8_t* __restrict__ jm, T* __restrict__ px, T* __restrict__ py,
                               T* __restrict__ pz) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = (int)blockIdx.z - 2;
    if (i0 >= g.nx || j0 >= g.ny) return;
    const long q = gidx(g, i0, j0, k);
    const unsigned b = jm[q];
    if (!(b & 1u)) px[q] = T(0.);
    if (!(b & 2u)) py[q] = T(0.);
    if (!(b & 4u)) pz[q] = T(0.);
    jm[q] = 7;
}

int launch_jump_densify(Ctx& c, bool old) {
    dim3 blk(BX, BY, 1);
    dim3 grd = grid3(c.g, c.g.nz + 6, blk);
    uint8_t* jm = old ? c.jmask_old : c.jmask;
    double *px = old ? c.p_xold : c.p_x, *py = old ? c.p_yold : c.p_y, *pz = old ? c.p_zold : c.p_z;
    if (c.fp32)
        hipLaunchKernelGGL(k_jump_densify<float>, grd, blk, 0, c.stream, c.g, jm, as<float>(px), as<float>(py),
                           as<float>(pz));
    else
        hipLaunchKernelGGL(k_jump_densify<double>, grd, blk, 0, c.stream, c.g, jm, px, py, pz);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

template <class T>
__global__ void k_fillps(Grid g, T coef, const T* __restrict__ du, const T* __restrict__ dv, const T* __restrict__ dw,
                         T* __restrict__ rhs) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    rhs[gidx(g, i0, j0, k)] = fillps_rhs<T>(g, coef, du, dv, dw, i0, j0, k);
}

int launch_fillps(Ctx& c) {
    ProfScope ps(c, "k_fillps");
    dim3 blk(BX, BY, 1);
    dim3 grd = grid3(c.g, c.g.nz, blk);
    if (c.fp32)
        hipLaunchKernelGGL(k_fillps<float>, grd, blk, 0, c.stream, c.g, (float)(c.rho0 / c.dt), as<float>(c.dudt),
                           as<float>(c.dvdt), as<float>(c.dwdt), as<float>(c.frhs));
    else
        hipLaunchKernelGGL(k_fillps<double>, grd, blk, 0, c.stream, c.g, c.rho0 / c.dt, (const double*)c.dudt,
                           (const double*)c.dvdt, (const double*)c.dwdt, c.frhs);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// rhs = rhs + p_jump (main.f90:142) over the interior
template <class T>
__global__ void k_add(Grid g, T* __restrict__ p, const T* __restrict__ a) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    long q = gidx(g, i0, j0, k);
    p[q] = p[q] + a[q];
}

int launch_add_pjump(Ctx& c) {
    ProfScope ps(c, "k_add");
    dim3 blk(BX, BY, 1);
    dim3 grd = grid3(c.g, c.g.nz, blk);
    if (c.fp32) hipLaunchKernelGGL(k_add<float>, grd, blk, 0, c.stream, c.g, as<float>(c.frhs), as<float>(c.p_jump));
    else hipLaunchKernelGGL(k_add<double>, grd, blk, 0, c.stream, c.g, c.frhs, (const double*)c.p_jump);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

}  // namespace icls
