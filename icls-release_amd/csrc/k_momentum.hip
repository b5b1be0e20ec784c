// k_momentum.hip -- ab2 + mom_main (convection, diffusion, buoyancy, AB2 and the
// split-pressure term), the velocity/pressure boundary conditions, correc, the
// pressure rotation and normalisation, cell-centre velocities and the diagnostic
// reductions (chkdiv, chkdt, max_vel_at_center).
#include "icls_internal.hpp"
#include "icls_reduce.hpp"
#include "icls_stencil.hpp"

namespace icls {

static constexpr int BX = 64, BY = 4;
static constexpr int KCH = 8;

struct MomParams {
    double dt, visc, iFr, g0, g1, g2, rho0, inv_rho0, rho1, forceytot;
    int rho_equal;
};

// ab2 (ab2.f90:49-129) with mom_main (mom.f90:13-62), CNTR2_conv (mom.f90:67-130),
// delta_diff (mom.f90:134-273), NS_grav (mom.f90:277-301) and momxp/yp/zp
// (mom.f90:779-851), one interior cell per thread.  Face densities are the
// arithmetic means of update_materials (misc.f90:151-172), formed here from rho.
template <int RHS_ONLY, bool P2>
__global__ void k_momentum(Grid g, MomParams m, const double* __restrict__ dscal, const double* __restrict__ u,
                           const double* __restrict__ v, const double* __restrict__ w, const double* __restrict__ miu,
                           const double* __restrict__ rho, double* __restrict__ duold, double* __restrict__ dvold,
                           double* __restrict__ dwold, const double* __restrict__ pnew, const double* __restrict__ pold,
                           const double* __restrict__ p_x, const double* __restrict__ p_y,
                           const double* __restrict__ p_z, const double* __restrict__ p_xold,
                           const double* __restrict__ p_yold, const double* __restrict__ p_zold,
                           double* __restrict__ dudt, double* __restrict__ dvdt, double* __restrict__ dwdt) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    const int im = wrap(i0 - 1, g.nx), ip = wrap(i0 + 1, g.nx), jm = wrap(j0 - 1, g.ny), jp = wrap(j0 + 1, g.ny);
    const int km = k - 1, kp = k + 1;
    auto I = [&](int a, int b, int c) { return gidx(g, a, b, c); };
    const Sp<double, P2> sx{g.dx, g.dxi}, sy{g.dy, g.dyi}, sz{g.dz, g.dzi};
    const int i = i0, j = j0;
    const long q = I(i, j, k);

    // ---- convection (CNTR2)
    double uuip = 0.25 * (u[I(ip, j, k)] + u[q]) * (u[I(ip, j, k)] + u[q]);
    double uuim = 0.25 * (u[I(im, j, k)] + u[q]) * (u[I(im, j, k)] + u[q]);
    double uvjp = 0.25 * (u[I(i, jp, k)] + u[q]) * (v[I(ip, j, k)] + v[q]);
    double uvjm = 0.25 * (u[I(i, jm, k)] + u[q]) * (v[I(ip, jm, k)] + v[I(i, jm, k)]);
    double uwkp = 0.25 * (u[I(i, j, kp)] + u[q]) * (w[I(ip, j, k)] + w[q]);
    double uwkm = 0.25 * (u[I(i, j, km)] + u[q]) * (w[I(ip, j, km)] + w[I(i, j, km)]);
    double RCu = sx.div(-uuip + uuim) + sy.div(-uvjp + uvjm) + sz.div(-uwkp + uwkm);
    double uvip = 0.25 * (u[q] + u[I(i, jp, k)]) * (v[q] + v[I(ip, j, k)]);
    double uvim = 0.25 * (u[I(im, j, k)] + u[I(im, jp, k)]) * (v[q] + v[I(im, j, k)]);
    double vvjp = 0.25 * (v[q] + v[I(i, jp, k)]) * (v[q] + v[I(i, jp, k)]);
    double vvjm = 0.25 * (v[q] + v[I(i, jm, k)]) * (v[q] + v[I(i, jm, k)]);
    double wvkp = 0.25 * (w[q] + w[I(i, jp, k)]) * (v[I(i, j, kp)] + v[q]);
    double wvkm = 0.25 * (w[I(i, j, km)] + w[I(i, jp, km)]) * (v[I(i, j, km)] + v[q]);
    double RCv = sx.div(-uvip + uvim) + sy.div(-vvjp + vvjm) + sz.div(-wvkp + wvkm);
    double uwip = 0.25 * (w[q] + w[I(ip, j, k)]) * (u[q] + u[I(i, j, kp)]);
    double uwim = 0.25 * (w[q] + w[I(im, j, k)]) * (u[I(im, j, k)] + u[I(im, j, kp)]);
    double vwjp = 0.25 * (w[q] + w[I(i, jp, k)]) * (v[q] + v[I(i, j, kp)]);
    double vwjm = 0.25 * (w[q] + w[I(i, jm, k)]) * (v[I(i, jm, k)] + v[I(i, jm, kp)]);
    double wwkp = 0.25 * (w[q] + w[I(i, j, kp)]) * (w[q] + w[I(i, j, kp)]);
    double wwkm = 0.25 * (w[q] + w[I(i, j, km)]) * (w[q] + w[I(i, j, km)]);
    double RCw = sx.div(-uwip + uwim) + sy.div(-vwjp + vwjm) + sz.div(-wwkp + wwkm);

    // ---- face densities
    double rq = rho[q];
    double rho_u, rho_v, rho_w;
    if (m.rho_equal) { rho_u = m.rho1; rho_v = m.rho1; rho_w = m.rho1; }
    else {
        rho_u = (rq + rho[I(ip, j, k)]) / 2.;
        rho_v = (rq + rho[I(i, jp, k)]) / 2.;
        rho_w = (rq + rho[I(i, j, kp)]) / 2.;
    }

    // ---- diffusion (delta formulation)
    const double mq = miu[q];
    double RDu, RDv, RDw;
    {
        double dudxp = sx.div(u[I(ip, j, k)] - u[q]);
        double dudxm = sx.div(u[q] - u[I(im, j, k)]);
        double dudyp = sy.div(u[I(i, jp, k)] - u[q]);
        double dudym = sy.div(u[q] - u[I(i, jm, k)]);
        double dvdxp = sx.div(v[I(ip, j, k)] - v[q]);
        double dvdxm = sx.div(v[I(ip, jm, k)] - v[I(i, jm, k)]);
        double dudzp = sz.div(u[I(i, j, kp)] - u[q]);
        double dudzm = sz.div(u[q] - u[I(i, j, km)]);
        double dwdxp = sx.div(w[I(ip, j, k)] - w[q]);
        double dwdxm = sx.div(w[I(ip, j, km)] - w[I(i, j, km)]);
        double ma1 = 0.25 * (mq + miu[I(i, jm, k)] + miu[I(ip, jm, k)] + miu[I(ip, j, k)]);
        double ma2 = 0.25 * (mq + miu[I(i, jp, k)] + miu[I(ip, jp, k)] + miu[I(ip, j, k)]);
        double ma3 = 0.25 * (mq + miu[I(i, j, km)] + miu[I(ip, j, km)] + miu[I(ip, j, k)]);
        double ma4 = 0.25 * (mq + miu[I(i, j, kp)] + miu[I(ip, j, kp)] + miu[I(ip, j, k)]);
        double RD1 = sx.div(2. * (dudxp * miu[I(ip, j, k)] - dudxm * mq));
        double RD2 = sy.div(dudyp * ma2 - dudym * ma1);
        double RD3 = sy.div(dvdxp * ma2 - dvdxm * ma1);
        double RD4 = sz.div(dudzp * ma4 - dudzm * ma3);
        double RD5 = sz.div(dwdxp * ma4 - dwdxm * ma3);
        RDu = (RD1 + RD2 + RD3 + RD4 + RD5) / rho_u;
    }
    {
        double dvdxp = sx.div(v[I(ip, j, k)] - v[q]);
        double dvdxm = sx.div(v[q] - v[I(im, j, k)]);
        double dudyp = sy.div(u[I(i, jp, k)] - u[q]);
        double dudym = sy.div(u[I(im, jp, k)] - u[I(im, j, k)]);
        double dvdyp = sy.div(v[I(i, jp, k)] - v[q]);
        double dvdym = sy.div(v[q] - v[I(i, jm, k)]);
        double dvdzp = sz.div(v[I(i, j, kp)] - v[q]);
        double dvdzm = sz.div(v[q] - v[I(i, j, km)]);
        double dwdyp = sy.div(w[I(i, jp, k)] - w[q]);
        double dwdym = sy.div(w[I(i, jp, km)] - w[I(i, j, km)]);
        double ma2 = 0.25 * (mq + miu[I(i, jp, k)] + miu[I(ip, jp, k)] + miu[I(ip, j, k)]);
        double ma5 = 0.25 * (mq + miu[I(i, jp, k)] + miu[I(im, jp, k)] + miu[I(im, j, k)]);
        double ma6 = 0.25 * (mq + miu[I(i, j, km)] + miu[I(i, jp, km)] + miu[I(i, jp, k)]);
        double ma7 = 0.25 * (mq + miu[I(i, j, kp)] + miu[I(i, jp, kp)] + miu[I(i, jp, k)]);
        double RD1 = sx.div(dvdxp * ma2 - dvdxm * ma5);
        double RD2 = sx.div(dudyp * ma2 - dudym * ma5);
        double RD3 = sy.div(2. * (dvdyp * miu[I(i, jp, k)] - dvdym * mq));
        double RD4 = sz.div(dvdzp * ma7 - dvdzm * ma6);
        double RD5 = sz.div(dwdyp * ma7 - dwdym * ma6);
        RDv = (RD1 + RD2 + RD3 + RD4 + RD5) / rho_v;
    }
    {
        double dwdxp = sx.div(w[I(ip, j, k)] - w[q]);
        double dwdxm = sx.div(w[q] - w[I(im, j, k)]);
        double dudzp = sz.div(u[I(i, j, kp)] - u[q]);
        double dudzm = sz.div(u[I(im, j, kp)] - u[I(im, j, k)]);
        double dwdyp = sy.div(w[I(i, jp, k)] - w[q]);
        double dwdym = sy.div(w[q] - w[I(i, jm, k)]);
        double dvdzp = sz.div(v[I(i, j, kp)] - v[q]);
        double dvdzm = sz.div(v[I(i, jm, kp)] - v[I(i, jm, k)]);
        double dwdzp = sz.div(w[I(i, j, kp)] - w[q]);
        double dwdzm = sz.div(w[q] - w[I(i, j, km)]);
        double ma4 = 0.25 * (mq + miu[I(i, j, kp)] + miu[I(ip, j, kp)] + miu[I(ip, j, k)]);
        double ma8 = 0.25 * (mq + miu[I(i, j, kp)] + miu[I(im, j, kp)] + miu[I(im, j, k)]);
        double ma7 = 0.25 * (mq + miu[I(i, j, kp)] + miu[I(i, jp, kp)] + miu[I(i, jp, k)]);
        double ma9 = 0.25 * (mq + miu[I(i, j, kp)] + miu[I(i, jm, kp)] + miu[I(i, jm, k)]);
        double RD1 = sx.div(dwdxp * ma4 - dwdxm * ma8);
        double RD2 = sx.div(dudzp * ma4 - dudzm * ma8);
        double RD3 = sy.div(dwdyp * ma7 - dwdym * ma9);
        double RD4 = sy.div(dvdzp * ma7 - dvdzm * ma9);
        double RD5 = sz.div(2. * (dwdzp * miu[I(i, j, kp)] - dwdzm * mq));
        RDw = (RD1 + RD2 + RD3 + RD4 + RD5) / rho_w;
    }

    // ---- buoyancy and RHS
    const double rho_avr = dscal[DS_RHO_AVR];
    double RGu = m.iFr * (1. - rho_avr / rho_u) * m.g0;
    double RGv = m.iFr * (1. - rho_avr / rho_v) * m.g1;
    double RGw = m.iFr * (1. - rho_avr / rho_w) * m.g2;
    double dun = RCu + RDu * m.visc + RGu;
    double dvn = RCv + RDv * m.visc + RGv;
    double dwn = RCw + RDw * m.visc + RGw;

    if (RHS_ONLY) {   // init.f90:90, mom_main into duold
        duold[q] = dun; dvold[q] = dvn; dwold[q] = dwn;
        return;
    }
    // ---- AB2 (ab2.f90:55-65)
    const double dt = m.dt;
    double us = u[q] + dt * (1.5 * dun - 0.5 * duold[q]);
    double vs = v[q] + dt * (1.5 * dvn - 0.5 * dvold[q]);
    double ws = w[q] + dt * (1.5 * dwn - 0.5 * dwold[q]);
    duold[q] = dun; dvold[q] = dvn; dwold[q] = dwn;

    // ---- split pressure (ab2.f90:79-101, mom.f90:793,818,843)
    double phq = 2. * pnew[q] - pold[q];
    double phx = 2. * pnew[I(ip, j, k)] - pold[I(ip, j, k)];
    double phy = 2. * pnew[I(i, jp, k)] - pold[I(i, jp, k)];
    double phz = 2. * pnew[I(i, j, kp)] - pold[I(i, j, kp)];
    double pjx = 2. * p_x[q] - p_xold[q];
    double pjy = 2. * p_y[q] - p_yold[q];
    double pjz = 2. * p_z[q] - p_zold[q];
    double mpx = -(1. / rho_u - m.inv_rho0) * g.dxi * (phx - pjx - phq);
    double mpy = -(1. / rho_v - m.inv_rho0) * g.dyi * (phy - pjy - phq);
    double mpz = -(1. / rho_w - m.inv_rho0) * g.dzi * (phz - pjz - phq);
    dudt[q] = us + dt * mpx;
    dvdt[q] = vs + dt * mpy - m.forceytot * dt / rho_v;
    dwdt[q] = ws + dt * mpz;
}

static int momentum_launch(Ctx& c, int rhs_only) {
    ProfScope ps(c, "k_momentum");
    MomParams m;
    m.dt = c.dt; m.visc = c.visc; m.iFr = c.P.iFr;
    m.g0 = c.P.g_unit[0]; m.g1 = c.P.g_unit[1]; m.g2 = c.P.g_unit[2];
    m.rho0 = c.rho0; m.inv_rho0 = 1. / c.rho0; m.rho1 = c.P.rho1;
    m.forceytot = c.P.constPoi ? -12. * c.visc / (c.P.lz * c.P.lz) : 0.;
    m.rho_equal = c.P.rho1 == c.P.rho2;
    dim3 blk(BX, BY, 1);
#define MOM(R, P)                                                                                                  \
    hipLaunchKernelGGL((k_momentum<R, P>), grid3(c.g, c.g.nz, blk), blk, 0, c.stream, c.g, m, (const double*)c.dscal,   \
                       (const double*)c.unew, (const double*)c.vnew, (const double*)c.wnew, (const double*)c.miu,   \
                       (const double*)c.rho, c.duold, c.dvold, c.dwold, (const double*)c.pnew, (const double*)c.pold, \
                       (const double*)c.p_x, (const double*)c.p_y, (const double*)c.p_z, (const double*)c.p_xold,   \
                       (const double*)c.p_yold, (const double*)c.p_zold, c.dudt, c.dvdt, c.dwdt)
    if (c.g.pow2) {
        if (rhs_only) MOM(1, true); else MOM(0, true);
    } else {
        if (rhs_only) MOM(1, false); else MOM(0, false);
    }
#undef MOM
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}
int launch_momentum(Ctx& c) { return momentum_launch(c, 0); }
int launch_momentum_rhs_only(Ctx& c) { return momentum_launch(c, 1); }

// mom_main alone into caller arrays (init.f90:90 starts AB2 with it)
__global__ void k_copy_interior(Grid g, const double* __restrict__ a, double* __restrict__ b) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    long q = gidx(g, i0, j0, k);
    b[q] = a[q];
}

// ---------------------------------------------------------------- boundary conditions
// bounduvw, BC_in_z = 'Dirichl' (bound.f90:26-47); x/y by wrap
__global__ void k_bc_uvw(Grid g, double vlo, double vup, double* __restrict__ u, double* __restrict__ v,
                         double* __restrict__ w, const double* __restrict__ flag) {
    if (flag && flag[0] == 0.) return;
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    if (i0 >= g.nx || j0 >= g.ny) return;
    const int kmax = g.nz, k1 = g.nz + 1;
    auto I = [&](int k) { return gidx(g, i0, j0, k); };
    u[I(0)] = -u[I(1)];
    u[I(k1)] = -u[I(kmax)];
    v[I(0)] = 2. * vlo - v[I(1)];
    v[I(k1)] = 2. * vup - v[I(kmax)];
    w[I(0)] = 0.;
    w[I(kmax)] = 0.;
    w[I(k1)] = w[I(kmax - 1)];
}

static void wall_speeds(const Ctx& c, double& vlo, double& vup) {
    if (c.P.ns_np) { vlo = 0.0; vup = 0.0; }
    else { vlo = -c.P.wall_speed; vup = c.P.wall_speed; }
}

int launch_bc_uvw_pred(Ctx& c, double* u, double* v, double* w, const double* flag) {
    ProfScope ps(c, "k_bc_uvw");
    double vlo, vup;
    wall_speeds(c, vlo, vup);
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_bc_uvw, grid3(c.g, 1, blk), blk, 0, c.stream, c.g, vlo, vup, u, v, w, flag);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}
int launch_bc_uvw(Ctx& c, double* u, double* v, double* w) { return launch_bc_uvw_pred(c, u, v, w, nullptr); }

// boundp z part: Neumann (bound.f90:193-198)
__global__ void k_bc_p(Grid g, double* __restrict__ p, const double* __restrict__ flag) {
    if (flag && flag[0] == 0.) return;
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    if (i0 >= g.nx || j0 >= g.ny) return;
    p[gidx(g, i0, j0, 0)] = p[gidx(g, i0, j0, 1)];
    p[gidx(g, i0, j0, g.nz + 1)] = p[gidx(g, i0, j0, g.nz)];
}

int launch_bc_p_pred(Ctx& c, double* p, const double* flag) {
    ProfScope ps(c, "k_bc_p");
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_bc_p, grid3(c.g, 1, blk), blk, 0, c.stream, c.g, p, flag);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}
int launch_bc_p(Ctx& c, double* p) { return launch_bc_p_pred(c, p, nullptr); }

// ---------------------------------------------------------------- correc
// correc.f90:33-35
__global__ void k_correc(Grid g, double dt, double rho0, const double* __restrict__ p, const double* __restrict__ px,
                         const double* __restrict__ py, const double* __restrict__ pz, const double* __restrict__ du,
                         const double* __restrict__ dv, const double* __restrict__ dw, double* __restrict__ u,
                         double* __restrict__ v, double* __restrict__ w) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    long q = gidx(g, i0, j0, k);
    double pq = p[q];
    u[q] = du[q] - dt * g.dxi * (p[gidx(g, wrap(i0 + 1, g.nx), j0, k)] - pq - px[q]) / rho0;
    v[q] = dv[q] - dt * g.dyi * (p[gidx(g, i0, wrap(j0 + 1, g.ny), k)] - pq - py[q]) / rho0;
    w[q] = dw[q] - dt * g.dzi * (p[gidx(g, i0, j0, k + 1)] - pq - pz[q]) / rho0;
}

int launch_correc(Ctx& c) {
    ProfScope ps(c, "k_correc");
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_correc, grid3(c.g, c.g.nz, blk), blk, 0, c.stream, c.g, c.dt, c.rho0, (const double*)c.p,
                       (const double*)c.p_x, (const double*)c.p_y, (const double*)c.p_z, (const double*)c.dudt,
                       (const double*)c.dvdt, (const double*)c.dwdt, c.unew, c.vnew, c.wnew);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- pressure rotation
// norm = sum(sum(pnew(:,:,1),2),1)/(itot*jtot) over the full C-halo plane, halos
// included (main.f90:172-175); partial sums over rows of the (i1+1) x (j1+1) plane
// one block per row j = 0 .. j1 of the plane, threads over i = 0 .. i1
__global__ void k_plane_sum(Grid g, const double* __restrict__ p, double* __restrict__ part) {
    __shared__ double smem[16];
    const int j0 = wrap((int)blockIdx.x - 1, g.ny);
    double acc = 0.;
    for (int i = threadIdx.x; i <= g.nx + 1; i += blockDim.x) acc += p[gidx(g, wrap(i - 1, g.nx), j0, 1)];
    double r = block_reduce(acc, OpSum(), smem);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ void k_norm_final(const double* __restrict__ part, int n, double denom, double* dscal) {
    __shared__ double smem[16];
    double v = 0.;
    for (int q = threadIdx.x; q < n; q += blockDim.x) v += part[q];
    double r = block_reduce(v, OpSum(), smem);
    if (threadIdx.x == 0) dscal[DS_NORM] = r / denom;
}

int launch_pressure_norm(Ctx& c) {
    ProfScope ps(c, "k_plane_sum");
    int nb = c.g.ny + 2;
    hipLaunchKernelGGL(k_plane_sum, dim3(nb), dim3(256), 0, c.stream, c.g, (const double*)c.p, c.dpart);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, c.stream, (const double*)c.dpart, nb,
                       1. * c.g.nx * c.g.ny, c.dscal);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// pold <- pnew, pnew <- p, both shifted by the norm (main.f90:164-183), over
// k = 0 .. k1.  Runs in place: pold[q] from pnew[q], then pnew[q] from p[q].
__global__ void k_pressure_rotate(Grid g, int first, const double* __restrict__ dscal, const double* __restrict__ p,
                                  double* __restrict__ pnew, double* __restrict__ pold) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i0 >= g.nx || j0 >= g.ny) return;
    long q = gidx(g, i0, j0, k);
    const double norm = dscal[DS_NORM], norm_old = dscal[DS_NORM_OLD];
    double po = pnew[q];
    pold[q] = first ? po - norm : po + norm_old - norm;
    pnew[q] = p[q] - norm;
}

__global__ void k_norm_shift(double* dscal) {
    if (threadIdx.x == 0) dscal[DS_NORM_OLD] = dscal[DS_NORM];
}

int launch_pressure_rotate(Ctx& c, int first_step) {
    ProfScope ps(c, "k_pressure_rotate");
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_pressure_rotate, grid3(c.g, c.g.nz + 2, blk), blk, 0, c.stream, c.g, first_step,
                       (const double*)c.dscal, (const double*)c.p, c.pnew, c.pold);
    hipLaunchKernelGGL(k_norm_shift, dim3(1), dim3(64), 0, c.stream, c.dscal);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- cell-centre velocity
// average_vel_at_center (misc.f90:587-604)
__global__ void k_avg_vel(Grid g, const double* __restrict__ u, const double* __restrict__ v,
                          const double* __restrict__ w, double* __restrict__ uc, double* __restrict__ vc,
                          double* __restrict__ wc) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    long q = gidx(g, i0, j0, k);
    uc[q] = (u[gidx(g, wrap(i0 - 1, g.nx), j0, k)] + u[q]) / 2.;
    vc[q] = (v[gidx(g, i0, wrap(j0 - 1, g.ny), k)] + v[q]) / 2.;
    wc[q] = (w[gidx(g, i0, j0, k - 1)] + w[q]) / 2.;
}

int launch_avg_vel(Ctx& c) {
    ProfScope ps(c, "k_avg_vel");
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_avg_vel, grid3(c.g, c.g.nz, blk), blk, 0, c.stream, c.g, (const double*)c.unew,
                       (const double*)c.vnew, (const double*)c.wnew, c.ucc, c.vcc, c.wcc);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- NaN and diagnostics
// isnan(vnew(1,1,1)) (main.f90:207-210) into a sticky flag
__global__ void k_nan_check(Grid g, const double* __restrict__ v, double* dscal) {
    if (threadIdx.x != 0) return;
    if (isnan(v[gidx(g, 0, 0, 1)])) dscal[DS_NAN_FLAG] = 1.;
}

int launch_nan_check(Ctx& c) {
    hipLaunchKernelGGL(k_nan_check, dim3(1), dim3(64), 0, c.stream, c.g, (const double*)c.vnew, c.dscal);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

static inline dim3 red_grid(const Grid& g, int nk) {
    return dim3((g.nx + BX - 1) / BX, (g.ny + BY - 1) / BY, (nk + KCH - 1) / KCH);
}
__device__ __forceinline__ int block_linear() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }

// chkdiv (chkdiv.f90:28-48): sum of div and max |div| with its first location in
// (k, j, i) loop order; partials [sum | max | location]
__global__ void k_chkdiv(Grid g, const double* __restrict__ u, const double* __restrict__ v,
                         const double* __restrict__ w, double* __restrict__ part, int nb) {
    __shared__ double smem[16];
    __shared__ double smax[256];
    __shared__ long sloc[256];
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    int tid = threadIdx.x + BX * threadIdx.y;
    double s = 0., mx = 0.;
    long loc = 0x7fffffffffffffffL;
    if (i0 < g.nx && j0 < g.ny)
        for (int kk = 0; kk < KCH; kk++) {
            int k = 1 + blockIdx.z * KCH + kk;
            if (k > g.nz) break;
            long q = gidx(g, i0, j0, k);
            double div = (w[q] - w[gidx(g, i0, j0, k - 1)]) * g.dzi + (v[q] - v[gidx(g, i0, wrap(j0 - 1, g.ny), k)]) * g.dyi +
                         (u[q] - u[gidx(g, wrap(i0 - 1, g.nx), j0, k)]) * g.dxi;
            s += div;
            double a = fabs(div);
            long lin = (long)i0 + (long)g.nx * ((long)j0 + (long)g.ny * (k - 1));
            if (a > mx || (a == mx && a > 0. && lin < loc)) { mx = a; loc = lin; }
        }
    double r = block_reduce(s, OpSum(), smem);
    smax[tid] = mx; sloc[tid] = loc;
    __syncthreads();
    if (tid == 0) {
        double bm = smax[0]; long bl = sloc[0];
        for (int t = 1; t < BX * BY; t++)
            if (smax[t] > bm || (smax[t] == bm && sloc[t] < bl)) { bm = smax[t]; bl = sloc[t]; }
        int b = block_linear();
        part[b] = r; part[nb + b] = bm; part[2 * nb + b] = (double)bl;
    }
}

__global__ void k_chkdiv_final(const double* __restrict__ part, int nb, double* dscal) {
    __shared__ double smem[16];
    double v = 0.;
    for (int q = threadIdx.x; q < nb; q += blockDim.x) v += part[q];
    double r = block_reduce(v, OpSum(), smem);
    if (threadIdx.x == 0) {
        double bm = 0., bl = 1e300;
        for (int b = 0; b < nb; b++) {
            double m = part[nb + b], l = part[2 * nb + b];
            if (m > bm || (m == bm && l < bl)) { bm = m; bl = l; }
        }
        dscal[DS_DIV_TOT] = r; dscal[DS_DIV_MAX] = bm; dscal[DS_DIV_MAX + 16] = bl;
    }
}

int launch_chkdiv(Ctx& c) {
    ProfScope ps(c, "k_chkdiv");
    dim3 blk(BX, BY, 1);
    dim3 grd = red_grid(c.g, c.g.nz);
    int nb = grd.x * grd.y * grd.z;
    hipLaunchKernelGGL(k_chkdiv, grd, blk, 0, c.stream, c.g, (const double*)c.unew, (const double*)c.vnew,
                       (const double*)c.wnew, c.dpart, nb);
    hipLaunchKernelGGL(k_chkdiv_final, dim3(1), dim3(256), 0, c.stream, (const double*)c.dpart, nb, c.dscal);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// chkdt (chkdt.f90:38-52): max |face-to-centre average| over 1..i1, 1..j1, 1..k1
__global__ void k_umax(Grid g, const double* __restrict__ u, const double* __restrict__ v,
                       const double* __restrict__ w, double* __restrict__ part, int nb) {
    __shared__ double smem[16];
    int i = blockIdx.x * BX + threadIdx.x + 1, j = blockIdx.y * BY + threadIdx.y + 1;   // Fortran 1 .. i1
    double mu = 0., mv = 0., mw = 0.;
    if (i <= g.nx + 1 && j <= g.ny + 1) {
        int i0 = wrap(i - 1, g.nx), j0 = wrap(j - 1, g.ny), im = wrap(i - 2, g.nx), jm = wrap(j - 2, g.ny);
        for (int kk = 0; kk < KCH; kk++) {
            int k = 1 + blockIdx.z * KCH + kk;
            if (k > g.nz + 1) break;
            long q = gidx(g, i0, j0, k);
            mu = fmax(mu, fabs((u[q] + u[gidx(g, im, j0, k)]) / 2.));
            mv = fmax(mv, fabs((v[q] + v[gidx(g, i0, jm, k)]) / 2.));
            mw = fmax(mw, fabs((w[q] + w[gidx(g, i0, j0, k - 1)]) / 2.));
        }
    }
    double a = block_reduce(mu, OpMax(), smem);
    double b = block_reduce(mv, OpMax(), smem);
    double c2 = block_reduce(mw, OpMax(), smem);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        int bl = block_linear();
        part[bl] = a; part[nb + bl] = b; part[2 * nb + bl] = c2;
    }
}

int launch_chkdt_umax(Ctx& c) {
    ProfScope ps(c, "k_umax");
    dim3 blk(BX, BY, 1);
    dim3 grd((c.g.nx + 1 + BX - 1) / BX, (c.g.ny + 1 + BY - 1) / BY, (c.g.nz + 1 + KCH - 1) / KCH);
    int nb = grd.x * grd.y * grd.z;
    hipLaunchKernelGGL(k_umax, grd, blk, 0, c.stream, c.g, (const double*)c.unew, (const double*)c.vnew,
                       (const double*)c.wnew, c.dpart, nb);
    for (int t = 0; t < 3; t++)
        hipLaunchKernelGGL(k_reduce_final<OpMax>, dim3(1), dim3(256), 0, c.stream, (const double*)(c.dpart + t * nb), nb,
                           0., c.dscal + DS_UMAX + t);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// max_vel_at_center (misc.f90:621-632)
__global__ void k_vmag_max(Grid g, const double* __restrict__ uc, const double* __restrict__ vc,
                           const double* __restrict__ wc, double* __restrict__ part) {
    __shared__ double smem[16];
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    double m = 0.;
    if (i0 < g.nx && j0 < g.ny)
        for (int kk = 0; kk < KCH; kk++) {
            int k = 1 + blockIdx.z * KCH + kk;
            if (k > g.nz) break;
            long q = gidx(g, i0, j0, k);
            m = fmax(m, sqrt(uc[q] * uc[q] + vc[q] * vc[q] + wc[q] * wc[q]));
        }
    double r = block_reduce(m, OpMax(), smem);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[block_linear()] = r;
}

int launch_max_vel(Ctx& c) {
    ProfScope ps(c, "k_vmag_max");
    dim3 blk(BX, BY, 1);
    dim3 grd = red_grid(c.g, c.g.nz);
    int nb = grd.x * grd.y * grd.z;
    hipLaunchKernelGGL(k_vmag_max, grd, blk, 0, c.stream, c.g, (const double*)c.ucc, (const double*)c.vcc,
                       (const double*)c.wcc, c.dpart);
    hipLaunchKernelGGL(k_reduce_final<OpMax>, dim3(1), dim3(256), 0, c.stream, (const double*)c.dpart, nb, 0.,
                       c.dscal + DS_VMAX);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- initial velocity
// init_vel_p 'cou' (init.f90:191-210): v = v_lower + shear_rate*(k-0.5)*dz, k = 0..k1
__global__ void k_init_couette(Grid g, double v_lower, double shear_rate, double* __restrict__ v) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i0 >= g.nx || j0 >= g.ny) return;
    double coorz = (k - 0.5) * g.dz;
    v[gidx(g, i0, j0, k)] = v_lower + shear_rate * coorz;
}

int launch_init_velocity(Ctx& c) {
    if (c.P.iniu != ICLS_INIU_COU) return ICLS_OK;
    double v_upper = c.P.wall_speed, v_lower = -c.P.wall_speed;
    double shear_rate = (v_upper - v_lower) / c.P.lz;
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_init_couette, grid3(c.g, c.g.nz + 2, blk), blk, 0, c.stream, c.g, v_lower, shear_rate, c.vnew);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// fill a whole padded field (used instead of memset so graph capture stays safe)
__global__ void k_fill(long n, double* __restrict__ f, double v) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) f[t] = v;
}

int launch_fill(Ctx& c, double* f, double v) {
    long n = c.nfield;
    int nb = (int)((n + 255) / 256);
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(256), 0, c.stream, n, f, v);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

int launch_copy_interior(Ctx& c, const double* a, double* b) {
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_copy_interior, grid3(c.g, c.g.nz, blk), blk, 0, c.stream, c.g, a, b);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

}  // namespace icls
