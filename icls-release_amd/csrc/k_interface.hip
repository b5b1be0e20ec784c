// k_interface.hip -- least-squares curvature (get_curvature), its smoothing
// (re_curvature), the ghost-fluid pressure jumps (gfm4p) and the Poisson right-hand
// side (fillps), fused where the reference applies them back to back.
#include "icls_internal.hpp"
#include "icls_stencil.hpp"

namespace icls {

static constexpr int BX = 64, BY = 4;

// 27-point gather in reshape order (i fastest, then j, then k), interface.f90:59
__device__ __forceinline__ void gather27(const Grid& g, const double* __restrict__ f, int i0, int j0, int k, double* v) {
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
__device__ __forceinline__ double lsq_row(const double* __restrict__ A, int r, const double* v) {
    double s = 0.;
#pragma unroll
    for (int q = 0; q < 27; q++) s = s + A[r + 10 * q] * v[q];
    return s;
}

// get_curvature (interface.f90:20-119): for |phi| <= NarrowBand_2 - dz fit the
// quadratic of init_matrix to the 27 values, kappa = -div(grad phi/|grad phi|) with
// the 0.9 gradient filter and the +-1/dz and 1e-9 limits; zero elsewhere.  The
// z-ghost layers of kappa are never written and stay zero (interface.f90:49).
__global__ void k_curvature(Grid g, double NB2, const double* __restrict__ A, const double* __restrict__ phi,
                            double* __restrict__ kappa, double* __restrict__ nm, const double* __restrict__ flag) {
    if (flag && flag[0] == 0.) return;
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    long q = gidx(g, i0, j0, k);
    const double dz = g.dz, max_kappa = 1. / dz;
    double cv = 0., m = 1.;
    if (fabs(phi[q]) <= NB2 - 1. * dz) {
        double v[27];
        gather27(g, phi, i0, j0, k, v);
        double dphidx = lsq_row(A, 1, v), dphidy = lsq_row(A, 2, v), dphidz = lsq_row(A, 3, v);
        double d2phidx = lsq_row(A, 4, v), d2phidy = lsq_row(A, 5, v), d2phidz = lsq_row(A, 6, v);
        double d2phidxdy = lsq_row(A, 7, v), d2phidxdz = lsq_row(A, 8, v), d2phidydz = lsq_row(A, 9, v);
        double norm_val = sqrt(dphidx * dphidx + dphidy * dphidy + dphidz * dphidz);
        m = norm_val;
        if (norm_val < 0.9) {
            cv = 0.;
        } else {
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
    }
    kappa[q] = cv;
    if (nm) nm[q] = m;
}

int launch_curvature(Ctx& c, const double* phi, double* kappa, double* normal_m, const double* flag) {
    ProfScope ps(c, "k_curvature");
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_curvature, grid3(c.g, c.g.nz, blk), blk, 0, c.stream, c.g, c.NB2, (const double*)c.pinvA, phi,
                       kappa, normal_m, flag);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// re_curvature (interface.f90:186-218): for |phi| < 2 dz the constant term of the
// least-squares fit to the 27 curvature values; zero elsewhere.
__global__ void k_recurvature(Grid g, const double* __restrict__ A, const double* __restrict__ phi,
                              const double* __restrict__ kappa, double* __restrict__ kappa2) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    long q = gidx(g, i0, j0, k);
    double r = 0.;
    if (fabs(phi[q]) < 2. * g.dz) {
        double v[27];
        gather27(g, kappa, i0, j0, k, v);
        r = lsq_row(A, 0, v);
    }
    kappa2[q] = r;
}

int launch_recurvature(Ctx& c, const double* phi, const double* kappa, double* kappa2) {
    ProfScope ps(c, "k_recurvature");
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_recurvature, grid3(c.g, c.g.nz, blk), blk, 0, c.stream, c.g, (const double*)c.pinvA, phi, kappa,
                       kappa2);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// fillps (fillps.f90:36-53): rho0/dt * div(u*)
__device__ __forceinline__ double fillps_rhs(const Grid& g, double coef, const double* __restrict__ du,
                                             const double* __restrict__ dv, const double* __restrict__ dw, int i0,
                                             int j0, int k) {
    long q = gidx(g, i0, j0, k);
    double vel_x = (du[q] - du[gidx(g, wrap(i0 - 1, g.nx), j0, k)]) * g.dxi;
    double vel_y = (dv[q] - dv[gidx(g, i0, wrap(j0 - 1, g.ny), k)]) * g.dyi;
    double vel_z = (dw[q] - dw[gidx(g, i0, j0, k - 1)]) * g.dzi;
    return coef * (vel_x + vel_y + vel_z);
}

// gfm4p jumps of one level set at one cell (gfm.f90:99-215).  Returns the separate
// jump Laplacian pj_sep and the + face jumps px/py/pz (already sign-flipped inside
// the drop); false when the cell is outside the |phi| <= 2 dz shell.
__device__ __forceinline__ bool gfm_cell(const Grid& g, double nondi, const double* __restrict__ ls,
                                         const double* __restrict__ cvf, int i0, int j0, int k, double& pjs, double& pxs,
                                         double& pys, double& pzs) {
    long q = gidx(g, i0, j0, k);
    double phc = ls[q];
    if (!(fabs(phc) <= 2. * g.dz)) return false;
    const double dx2 = g.dx * g.dx, dy2 = g.dy * g.dy, dz2 = g.dz * g.dz;
    double ph[3], cv[3], jump[3];
    double face[3] = {0., 0., 0.};
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        long qm, qp;
        if (ax == 0) { qm = gidx(g, wrap(i0 - 1, g.nx), j0, k); qp = gidx(g, wrap(i0 + 1, g.nx), j0, k); }
        else if (ax == 1) { qm = gidx(g, i0, wrap(j0 - 1, g.ny), k); qp = gidx(g, i0, wrap(j0 + 1, g.ny), k); }
        else { qm = gidx(g, i0, j0, k - 1); qp = gidx(g, i0, j0, k + 1); }
        ph[0] = ls[qm]; ph[1] = phc; ph[2] = ls[qp];
        cv[0] = cvf[qm]; cv[1] = cvf[q]; cv[2] = cvf[qp];
        double jm = 0., jp = 0.;
        if (ph[1] * ph[0] < 0.) jm = nondi * (-icurvature<double>(-1, ph + 1, cv + 1));
        if (ph[1] * ph[2] < 0.) { jp = nondi * (-icurvature<double>(1, ph + 1, cv + 1)); face[ax] = jp; }
        jump[ax] = jm + 0. + jp;   // sum(jp_x) over (-1, 0, 1)
    }
    if (phc > 0.) {
        pjs = (jump[0] / dx2 + jump[1] / dy2 + jump[2] / dz2);
        pxs = face[0]; pys = face[1]; pzs = face[2];
    } else if (phc < 0.) {
        pjs = -(jump[0] / dx2 + jump[1] / dy2 + jump[2] / dz2);
        pxs = -face[0]; pys = -face[1]; pzs = -face[2];
    } else {
        pjs = 0.; pxs = 0.; pys = 0.; pzs = 0.;
    }
    return true;
}

// MODE 0: first level set -- write p_jump, p_x, p_y, p_z over the interior (zero off
//         the shell), gfm.f90:71-79 and 219-224.
// MODE 1: further level sets -- add inside their shell (gfm.f90:226-232).
// MODE 2: single level set fused with fillps and main.f90:142 -- p = rhs + p_jump.
template <int MODE>
__global__ void k_gfm(Grid g, double nondi, double coef, const double* __restrict__ ls, const double* __restrict__ cvf,
                      double* __restrict__ pjump, double* __restrict__ px, double* __restrict__ py,
                      double* __restrict__ pz, const double* __restrict__ du, const double* __restrict__ dv,
                      const double* __restrict__ dw, double* __restrict__ p) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    long q = gidx(g, i0, j0, k);
    double pjs = 0., pxs = 0., pys = 0., pzs = 0.;
    bool shell = gfm_cell(g, nondi, ls, cvf, i0, j0, k, pjs, pxs, pys, pzs);
    if (MODE == 1) {
        if (shell) {
            pjump[q] = pjump[q] + pjs;
            px[q] = px[q] + pxs; py[q] = py[q] + pys; pz[q] = pz[q] + pzs;
        }
        return;
    }
    px[q] = pxs; py[q] = pys; pz[q] = pzs;
    if (pjump) pjump[q] = pjs;
    if (MODE == 2) p[q] = fillps_rhs(g, coef, du, dv, dw, i0, j0, k) + pjs;
}

int launch_gfm(Ctx& c, int l, const double* kappa2, bool fuse_fillps, bool store_pjump) {
    ProfScope ps(c, "k_gfm");
    dim3 blk(BX, BY, 1);
    dim3 grd = grid3(c.g, c.g.nz, blk);
    double nondi = 1. / c.P.We, coef = c.rho0 / c.dt;
    double* pj = store_pjump ? c.p_jump : nullptr;
    if (fuse_fillps)
        hipLaunchKernelGGL(k_gfm<2>, grd, blk, 0, c.stream, c.g, nondi, coef, (const double*)c.lvset[l], kappa2, pj, c.p_x,
                           c.p_y, c.p_z, (const double*)c.dudt, (const double*)c.dvdt, (const double*)c.dwdt, c.p);
    else if (l == 0)
        hipLaunchKernelGGL(k_gfm<0>, grd, blk, 0, c.stream, c.g, nondi, coef, (const double*)c.lvset[l], kappa2, c.p_jump,
                           c.p_x, c.p_y, c.p_z, (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                           (double*)nullptr);
    else
        hipLaunchKernelGGL(k_gfm<1>, grd, blk, 0, c.stream, c.g, nondi, coef, (const double*)c.lvset[l], kappa2, c.p_jump,
                           c.p_x, c.p_y, c.p_z, (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                           (double*)nullptr);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

__global__ void k_fillps(Grid g, double coef, const double* __restrict__ du, const double* __restrict__ dv,
                         const double* __restrict__ dw, double* __restrict__ p) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    p[gidx(g, i0, j0, k)] = fillps_rhs(g, coef, du, dv, dw, i0, j0, k);
}

int launch_fillps(Ctx& c) {
    ProfScope ps(c, "k_fillps");
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_fillps, grid3(c.g, c.g.nz, blk), blk, 0, c.stream, c.g, c.rho0 / c.dt, (const double*)c.dudt,
                       (const double*)c.dvdt, (const double*)c.dwdt, c.p);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// p = p + p_jump (main.f90:142) over the interior
__global__ void k_add(Grid g, double* __restrict__ p, const double* __restrict__ a) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    long q = gidx(g, i0, j0, k);
    p[q] = p[q] + a[q];
}

int launch_add_pjump(Ctx& c) {
    ProfScope ps(c, "k_add");
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_add, grid3(c.g, c.g.nz, blk), blk, 0, c.stream, c.g, c.p, (const double*)c.p_jump);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

}  // namespace icls
