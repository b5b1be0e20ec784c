// k_levelset.hip -- narrow-band level-set kernels: tube construction, SSP-RK3
// advection stages (HOUC5 / WENO5), WENO5 re-initialisation stages, z-wall
// extrapolation, initial sphere SDFs.
//
// Every kernel sweeps the whole padded field with one thread per cell; x and y wrap
// periodically, which replaces the 3-cell halo exchange updthalos_ls
// (bound.f90:760-797).  The narrow band is a byte mask over the interior instead of
// the reference's compacted index lists (levelset.f90:160-188); a dense masked sweep
// touches exactly the cells RK_nb updates (levelset.f90:353-361).
#include "icls_internal.hpp"
#include "icls_reduce.hpp"
#include "icls_stencil.hpp"

namespace icls {

static constexpr int BX = 64, BY = 4;

// ---------------------------------------------------------------- tube_construction
// Clamp phi to +-NarrowBand_2 over the whole array (levelset.f90:107-119) and mark
// the band: a cell is in the band unless all 13 points of its +-2 axis scans sit at
// |phi| == NarrowBand_2 (levelset.f90:125-156; a sum of non-positive terms is zero
// only when every term is).
__global__ void k_tube(Grid g, double NB2, const double* __restrict__ in, double* __restrict__ out,
                       uint8_t* __restrict__ mask) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, ks = blockIdx.z;
    if (i0 >= g.nx || j0 >= g.ny) return;
    int k = ks - 2;
    long q = gidx(g, i0, j0, k);
    double v = in[q];
    double c = v > NB2 ? NB2 : (v < -NB2 ? -NB2 : v);
    out[q] = c;
    if (k < 1 || k > g.nz) return;
    auto inside = [&](double x) { double y = x > NB2 ? NB2 : (x < -NB2 ? -NB2 : x); return fabs(y) < NB2; };
    bool m = fabs(c) < NB2;
    if (!m) {
        for (int d = -2; d <= 2 && !m; d++) {
            if (d == 0) continue;
            m = inside(in[gidx(g, wrap(i0 + d, g.nx), j0, k)]) || inside(in[gidx(g, i0, wrap(j0 + d, g.ny), k)]) ||
                inside(in[gidx(g, i0, j0, k + d)]);
        }
    }
    mask[q] = m ? 1 : 0;
}

int launch_tube(Ctx& c, const double* phi_in, double* phi_out) {
    ProfScope ps(c, "k_tube");
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_tube, grid3(c.g, c.g.nz + 6, blk), blk, 0, c.stream, c.g, c.NB2, phi_in, phi_out, c.mask);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- boundsls z part
// 'line' extrapolation of one ghost layer at each wall (bound.f90:404-411).  When
// res_ref is given, also folds |new - res_ref| of the two rewritten layers into the
// re-initialisation residual (levelset.f90:831).
__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v) {
    atomicMax((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}

// wave-wide max of a non-negative value followed by one atomic per wavefront; every
// lane of the wave must reach this call (out-of-range lanes pass 0)
__device__ __forceinline__ void wave_atomic_max(double* addr, double d) {
    for (int off = 32; off > 0; off >>= 1) d = fmax(d, __shfl_xor(d, off));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(addr, d);
}

__global__ void k_bc_ls(Grid g, double* __restrict__ phi, const double* __restrict__ res_ref, double* res,
                        const double* __restrict__ flag) {
    if (flag && flag[0] == 0.) return;
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    bool valid = i0 < g.nx && j0 < g.ny;
    double r = 0.;
    if (valid) {
        const int km = g.nz;
        double lo = 2. * phi[gidx(g, i0, j0, 1)] - phi[gidx(g, i0, j0, 2)];
        double hi = 2. * phi[gidx(g, i0, j0, km)] - phi[gidx(g, i0, j0, km - 1)];
        phi[gidx(g, i0, j0, 0)] = lo;
        phi[gidx(g, i0, j0, km + 1)] = hi;
        // res_ref: the two stashed wall layers of phi0, plane 0 then plane k1
        if (res_ref) r = fmax(fabs(lo - res_ref[i0 + g.sy * j0]), fabs(hi - res_ref[g.sz + i0 + g.sy * j0]));
    }
    if (res_ref) wave_atomic_max(res, r);
}

int launch_bc_ls_pred(Ctx& c, double* phi, const double* flag) {
    ProfScope ps(c, "k_bc_ls");
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_bc_ls, grid3(c.g, 1, blk), blk, 0, c.stream, c.g, phi, (const double*)nullptr, (double*)nullptr,
                       flag);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

int launch_bc_ls(Ctx& c, double* phi) {
    ProfScope ps(c, "k_bc_ls");
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_bc_ls, grid3(c.g, 1, blk), blk, 0, c.stream, c.g, phi, (const double*)nullptr, (double*)nullptr,
                       (const double*)nullptr);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// boundmls: three extrapolated layers per wall (bound.f90:539-549), init only
__global__ void k_boundmls(Grid g, double* __restrict__ y) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    if (i0 >= g.nx || j0 >= g.ny) return;
    const int kmax = g.nz, k1 = g.nz + 1;
    auto Y = [&](int k) -> double& { return y[gidx(g, i0, j0, k)]; };
    Y(0) = 2. * Y(1) - Y(2);
    Y(-1) = 2. * Y(0) - Y(1);
    Y(-2) = 2. * Y(-1) - Y(0);
    Y(k1) = 2. * Y(kmax) - Y(kmax - 1);
    Y(k1 + 1) = 2. * Y(k1) - Y(kmax);
    Y(k1 + 2) = 2. * Y(k1 + 1) - Y(k1);
}

int launch_boundmls(Ctx& c, double* phi) {
    ProfScope ps(c, "k_boundmls");
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_boundmls, grid3(c.g, 1, blk), blk, 0, c.stream, c.g, phi);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- advection stage
// dfdt of HOUC5_nb (hyperPDE.f90:397-495) or WENO5_nb (hyperPDE.f90:503-800) at one
// band cell.  The z stencil drops to HOUC3 at k = 2, kmax-1 and to first order at
// k = 1, kmax in both schemes.
template <int SCHEME, bool P2>
__device__ __forceinline__ double adv_rhs(const Grid& g, const double* __restrict__ phi, int i0, int j0, int k,
                                          double vu, double vv, double vw) {
    double f[7];
    const Sp<double, P2> sx{g.dx, g.dxi}, sy{g.dy, g.dyi}, sz{g.dz, g.dzi};
#pragma unroll
    for (int d = -3; d <= 3; d++) f[d + 3] = phi[gidx(g, wrap(i0 + d, g.nx), j0, k)];
    double udfdx, vdfdy, wdfdz;
    if (SCHEME == ICLS_ADV_HOUC5) {
        udfdx = sx.div(-vu * houc5<double>(vu, f + 3));
    } else {
        double dfx = (vu >= 0.) ? weno_minus(f + 3, sx) : weno_plus(f + 3, sx);
        udfdx = -vu * dfx;
    }
#pragma unroll
    for (int d = -3; d <= 3; d++) f[d + 3] = phi[gidx(g, i0, wrap(j0 + d, g.ny), k)];
    if (SCHEME == ICLS_ADV_HOUC5) {
        vdfdy = sy.div(-vv * houc5<double>(vv, f + 3));
    } else {
        // y differences are divided by dx in the reference (hyperPDE.f90:622-626)
        double dfy = (vv >= 0.) ? weno_minus(f + 3, sx) : weno_plus(f + 3, sx);
        vdfdy = -vv * dfy;
    }
    const int kmax = g.nz;
    if (k > 2 && k < kmax - 1) {
#pragma unroll
        for (int d = -3; d <= 3; d++) f[d + 3] = phi[gidx(g, i0, j0, k + d)];
        if (SCHEME == ICLS_ADV_HOUC5) {
            wdfdz = sz.div(-vw * houc5<double>(vw, f + 3));
        } else {
            double dfz = (vw >= 0.) ? weno_minus(f + 3, sz) : weno_plus(f + 3, sz);
            wdfdz = -vw * dfz;
        }
    } else {
#pragma unroll
        for (int d = -2; d <= 2; d++) f[d + 3] = phi[gidx(g, i0, j0, k + d)];
        double dfz = (k == 2 || k == kmax - 1) ? houc3<double>(vw, f + 3) : upwind1<double>(vw, f + 3);
        wdfdz = sz.div(-vw * dfz);
    }
    return udfdx + vdfdy + wdfdz;
}

// One SSP-RK3 stage of ls_advection + RK_nb (levelset.f90:235-243, 342-367):
//   stage 1: out = phi0 + dt*L(phi0)                    (band), phi0 elsewhere
//   stage 2: out = [3/4 phi0 + 1/4 phi1] + 1/4 dt L(phi1)
//   stage 3: out = [1/3 phi0 + 2/3 phi2] + 2/3 dt L(phi2)
// The bracket is the reference's whole-array temporary, applied to every stored
// cell including the z ghost layers.  flag (optional) predicates the launch on a
// device-side switch (used by the mass correction).
// VELF: vu/vv/vw are the face velocities unew/vnew/wnew and the cell-centre values
// are averaged here as average_vel_at_center does (misc.f90:595-597); otherwise they
// are cell-centre arrays already.
template <int SCHEME, int STAGE, bool P2, bool VELF>
__global__ void k_adv_stage(Grid g, double dt, const double* phi0, const double* __restrict__ phis, double* out,
                            const uint8_t* __restrict__ mask, const double* __restrict__ vu,
                            const double* __restrict__ vv, const double* __restrict__ vw, const double* __restrict__ flag) {
    if (flag && flag[0] == 0.) return;
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, ks = blockIdx.z;
    if (i0 >= g.nx || j0 >= g.ny) return;
    int k = ks - 2;
    long q = gidx(g, i0, j0, k);
    double comb, factor;
    if (STAGE == 1) { comb = phi0[q]; factor = 1.; }
    else if (STAGE == 2) { comb = 3. / 4. * phi0[q] + 1. / 4. * phis[q]; factor = 1. / 4.; }
    else { comb = 1. / 3. * phi0[q] + 2. / 3. * phis[q]; factor = 2. / 3.; }
    double r = comb;
    if (k >= 1 && k <= g.nz && mask[q]) {
        double uc, vc, wc;
        if (VELF) {
            uc = (vu[gidx(g, wrap(i0 - 1, g.nx), j0, k)] + vu[q]) / 2.;
            vc = (vv[gidx(g, i0, wrap(j0 - 1, g.ny), k)] + vv[q]) / 2.;
            wc = (vw[gidx(g, i0, j0, k - 1)] + vw[q]) / 2.;
        } else {
            uc = vu[q]; vc = vv[q]; wc = vw[q];
        }
        double dfdt = adv_rhs<SCHEME, P2>(g, phis, i0, j0, k, uc, vc, wc);
        r = comb + factor * dt * dfdt;
    }
    out[q] = r;
}

int launch_adv_stage(Ctx& c, int scheme, int stage, const double* phi0, const double* phis, double* out,
                     const double* vu, const double* vv, const double* vw, int velf, const double* /*kappa*/,
                     const double* flag) {
    dim3 blk(BX, BY, 1);
    dim3 grd = grid3(c.g, c.g.nz + 6, blk);
#define ADV1(S, T, P, V) \
    hipLaunchKernelGGL((k_adv_stage<S, T, P, V>), grd, blk, 0, c.stream, c.g, c.dt, phi0, phis, out, c.mask, vu, vv, vw, flag)
#define ADV(S, T)                                                          \
    if (scheme == S && stage == T) {                                       \
        ProfScope ps(c, T == 1 ? "k_adv_stage1" : "k_adv_stage23");        \
        if (c.g.pow2) {                                                    \
            if (velf) ADV1(S, T, true, true); else ADV1(S, T, true, false);   \
        } else {                                                           \
            if (velf) ADV1(S, T, false, true); else ADV1(S, T, false, false); \
        }                                                                  \
    }
    ADV(ICLS_ADV_HOUC5, 1) ADV(ICLS_ADV_HOUC5, 2) ADV(ICLS_ADV_HOUC5, 3)
    ADV(ICLS_ADV_WENO5, 1) ADV(ICLS_ADV_WENO5, 2) ADV(ICLS_ADV_WENO5, 3)
#undef ADV
#undef ADV1
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- re-initialisation
// Six one-sided WENO5 derivatives and the Godunov-type selection of WENO5_reinit
// (hyperPDE.f90:2662-2736, 2932-3197), evaluated where |phi| < NarrowBand_2 + dz.
// Along an axis whose seven values are equal (the clamped far field) both WENO5
// derivatives are exactly +0 (every difference is +0, every smoothness indicator 0,
// the weighted sum of zeros +0), so they are set without evaluating the weights.
template <bool P2>
__device__ __forceinline__ double reinit_rhs(const Grid& g, const double* __restrict__ phi, int i0, int j0, int k,
                                             double s) {
    double f[7];
    const Sp<double, P2> sx{g.dx, g.dxi}, sz{g.dz, g.dzi};
#pragma unroll
    for (int d = -3; d <= 3; d++) f[d + 3] = phi[gidx(g, wrap(i0 + d, g.nx), j0, k)];
    double A = 0., B = 0.;
    if (!flat7(f)) { A = weno_minus(f + 3, sx); B = weno_plus(f + 3, sx); }
#pragma unroll
    for (int d = -3; d <= 3; d++) f[d + 3] = phi[gidx(g, i0, wrap(j0 + d, g.ny), k)];
    double C = 0., D = 0.;
    if (!flat7(f)) { C = weno_minus(f + 3, sx); D = weno_plus(f + 3, sx); }
    double E = 0., F = 0.;
    if (k >= 3 && k <= g.nz - 2) {
#pragma unroll
        for (int d = -3; d <= 3; d++) f[d + 3] = phi[gidx(g, i0, j0, k + d)];
        if (!flat7(f)) { E = weno_minus(f + 3, sz); F = weno_plus(f + 3, sz); }
    } else {
        double pc = phi[gidx(g, i0, j0, k)];
        E = sz.div(pc - phi[gidx(g, i0, j0, k - 1)]);
        F = sz.div(phi[gidx(g, i0, j0, k + 1)] - pc);
    }
    double dydx, dydy, dydz;
    if ((B * s < 0.) && ((A + B) * s < 0.)) dydx = B;
    else if ((A * s > 0.) && ((A + B) * s > 0.)) dydx = A;
    else dydx = 0.5 * (A + B);
    if ((D * s < 0.) && ((C + D) * s < 0.)) dydy = D;
    else if ((C * s > 0.) && ((C + D) * s > 0.)) dydy = C;
    else dydy = 0.5 * (C + D);
    if ((F * s < 0.) && ((E + F) * s < 0.)) dydz = F;
    else if ((E * s > 0.) && ((E + F) * s > 0.)) dydz = E;
    else dydz = 0.5 * (E + F);
    return s * (1. - sqrt(dydx * dydx + dydy * dydy + dydz * dydz));
}

// One RK3 stage of ls_reinit (levelset.f90:798-810) with the smoothed sign of
// levelset.f90:668-693 recomputed from phi0 on the fly:
//   stage 1: out = phi0 + dtau*RHS(phi0)
//   stage 2: out = 3/4 phi0 + 1/4 (phi_n1 + dtau*RHS(phi_n1))
//   stage 3: out = 1/3 phi0 + 2/3 (phi_n2 + dtau*RHS(phi_n2))
// RHS is zero outside the interior band (WENO5_reinit zeroes it, hyperPDE.f90:2662).
// Stage 3 also folds |out - phi0| of every layer that boundsls does not rewrite into
// the residual (levelset.f90:831); stage 1 resets it.
// Each thread marches RKZ consecutive stored layers of one (i,j) column; the stage-3
// residual is folded per block and published with one atomic max per block (max is
// order independent, so the result does not depend on the schedule).
static constexpr int RKZ = 8;

template <int STAGE, bool P2>
__global__ void k_reinit_stage(Grid g, double NB2, double* res, const double* phi0, const double* __restrict__ phis,
                               double* out, double* __restrict__ stash) {
    __shared__ double smem[16];
    const int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    if (STAGE == 1 && i0 == 0 && j0 == 0 && blockIdx.z == 0) res[0] = 0.;
    const bool valid = i0 < g.nx && j0 < g.ny;
    const double dz = g.dz, dtau = 0.5 * dz;
    double dmax = 0.;
    for (int kk = 0; kk < RKZ; kk++) {
        const int ks = blockIdx.z * RKZ + kk;
        if (!valid || ks >= g.nz + 6) break;
        const int k = ks - 2;
        const long q = gidx(g, i0, j0, k);
        double rhs = 0.;
        if (k >= 1 && k <= g.nz) {
            double ps = phis[q];
            if (fabs(ps) < NB2 + dz) {
                double p0 = phi0[q], s;
                if (p0 < -dz) s = -1.;
                else if (p0 > dz) s = 1.;
                else s = p0 / sqrt(p0 * p0 + dz * dz);
                rhs = reinit_rhs<P2>(g, phis, i0, j0, k, s);
            }
        }
        double r;
        if (STAGE == 1) r = phi0[q] + dtau * rhs;
        else if (STAGE == 2) r = 3. / 4. * phi0[q] + 1. / 4. * (phis[q] + dtau * rhs);
        else r = 1. / 3. * phi0[q] + 2. / 3. * (phis[q] + dtau * rhs);
        if (STAGE == 3) {
            // stage 3 may run in place (out == phi0): keep phi0 of the two wall layers
            // that boundsls rewrites for the residual check in k_bc_ls
            double p0 = phi0[q];
            if (k == 0) stash[i0 + g.sy * j0] = p0;
            else if (k == g.nz + 1) stash[g.sz + i0 + g.sy * j0] = p0;
            else dmax = fmax(dmax, fabs(r - p0));
        }
        out[q] = r;
    }
    if (STAGE == 3) {
        double b = block_reduce(dmax, OpMax(), smem);
        if (threadIdx.x == 0 && threadIdx.y == 0) atomic_max_nonneg(res, b);
    }
}

int launch_reinit_stage(Ctx& c, int stage, const double* phi0, const double* phis, double* out) {
    ProfScope ps(c, stage == 1 ? "k_reinit_stage1" : "k_reinit_stage23");
    dim3 blk(BX, BY, 1);
    dim3 grd = grid3(c.g, (c.g.nz + 6 + RKZ - 1) / RKZ, blk);
    double* res = c.dscal + DS_REINIT_RES;
#define RS(S, P) \
    hipLaunchKernelGGL((k_reinit_stage<S, P>), grd, blk, 0, c.stream, c.g, c.NB2, res, phi0, phis, out, c.stash)
    if (c.g.pow2) {
        if (stage == 1) RS(1, true); else if (stage == 2) RS(2, true); else RS(3, true);
    } else {
        if (stage == 1) RS(1, false); else if (stage == 2) RS(2, false); else RS(3, false);
    }
#undef RS
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// boundsls after the last re-init stage, folding the two wall layers into the residual
// against the phi0 layers stage 3 stashed
int launch_reinit_residual(Ctx& c, const double* /*phi0*/, const double* phinew) {
    ProfScope ps(c, "k_bc_ls");
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_bc_ls, grid3(c.g, 1, blk), blk, 0, c.stream, c.g, (double*)phinew, (const double*)c.stash,
                       c.dscal + DS_REINIT_RES, (const double*)nullptr);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- initial level sets
// Sphere SDF with the +lx and +ly periodic images (init.f90:945-980; 'drp' is the
// lmax = 1 case, init.f90:717-753), evaluated at every stored cell.
__global__ void k_init_ls(Grid g, double lx, double ly, double xc, double yc, double zc, double R, double* __restrict__ ph) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, ks = blockIdx.z;
    if (i0 >= g.nx || j0 >= g.ny) return;
    int k = ks - 2, i = i0 + 1, j = j0 + 1;
    double coorz = (k - 0.5) / g.dzi;
    double coory = 0. + (j - 0.5) / g.dyi;
    double coorx = 0. + (i - 0.5) / g.dxi;
    double v = sqrt((coorx - xc) * (coorx - xc) + (coory - yc) * (coory - yc) + (coorz - zc) * (coorz - zc)) - R;
    double p1 = sqrt((coorx - xc) * (coorx - xc) + (coory - yc - ly) * (coory - yc - ly) + (coorz - zc) * (coorz - zc)) - R;
    if (v > 0.) v = fmin(v, p1);
    double p2 = sqrt((coorx - xc - lx) * (coorx - xc - lx) + (coory - yc) * (coory - yc) + (coorz - zc) * (coorz - zc)) - R;
    if (v > 0.) v = fmin(v, p2);
    ph[gidx(g, i0, j0, k)] = v;
}

int launch_init_levelset(Ctx& c, int l, double* phi) {
    ProfScope ps(c, "k_init_ls");
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_init_ls, grid3(c.g, c.g.nz + 6, blk), blk, 0, c.stream, c.g, c.P.lx, c.P.ly, c.P.drop_xc[l],
                       c.P.drop_yc[l], c.P.drop_zc[l], c.P.drop_r[l], phi);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

}  // namespace icls
