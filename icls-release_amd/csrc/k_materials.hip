// k_materials.hip -- update_materials (Heaviside, viscosity, density, volume), the
// mass-correction "inflation" sums and velocities, scalar bookkeeping kernels and the
// generic interior sums.
//
// T is the element type of the field arrays.  Every sum (volumes, interface area) is
// accumulated in double whatever T is.
#include "icls_internal.hpp"
#include "icls_reduce.hpp"
#include "icls_stencil.hpp"

namespace icls {

static constexpr int BX = 64, BY = 4;
static constexpr int KCH = 8;   // z layers folded per block in the reductions

// number of blocks of a (BX x BY) x (z chunks of KCH) sweep over k = k_lo .. k_hi
static inline dim3 red_grid(const Grid& g, int k_lo, int k_hi) {
    int nk = k_hi - k_lo + 1;
    return dim3((g.nx + BX - 1) / BX, (g.ny + BY - 1) / BY, (nk + KCH - 1) / KCH);
}
static inline int red_nblocks(dim3 d) { return (int)(d.x * d.y * d.z); }

// ---------------------------------------------------------------- update_materials
// misc.f90:100-147 for every cell of the C-halo range (x/y by wrap, z = 0..k1):
//   H_l  = Heaviside('0','trig',l,1.5 dz),  H = min_l H_l   (jiemian and h_rho)
//   miu  = miu1 H + miu2 (1 - H),  rho = rho1 H + rho2 (1 - H)
// and the partial interior sums of every H_l for volume(l) (misc.f90:128-134).
// With level-set windows (win: ox, oy, oz, wx, wy, wz per level set) a cell outside
// window l -- and its three ghost layers in z -- takes H_l = 1 without reading phi_l:
// the level set is in its far field there (phi_l >= NarrowBand_2 > eps), where the
// Heaviside is exactly 1.
__device__ __forceinline__ bool in_win(const int* __restrict__ w, int i0, int j0, int k) {
    return (unsigned)(i0 - w[0]) < (unsigned)w[3] && (unsigned)(j0 - w[1]) < (unsigned)w[4] && k >= w[2] - 2 &&
           k <= w[2] + w[5] + 3;
}

template <class T>
__device__ __forceinline__ T heav_l(const T* const* __restrict__ ls, const int* __restrict__ win, int l, long q,
                                    int i0, int j0, int k, T eps, T pi) {
    if (win && !in_win(win + 6 * l, i0, j0, k)) return T(1.);
    return heaviside_trig<T>(ls[l][q], eps, pi);
}

template <class T>
__global__ void k_materials(Grid g, int lmax, const T* const* __restrict__ ls, const int* __restrict__ win, T eps,
                            T pi, T miu1, T miu2, T rho1, T rho2, T* __restrict__ miu, T* __restrict__ rho,
                            T* __restrict__ jiemian, double* __restrict__ part, int nblocks) {
    __shared__ double smem[16];
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    bool valid = i0 < g.nx && j0 < g.ny;
    // up to 4 level sets summed in registers per pass; more are folded pass by pass
    for (int l0 = 0; l0 < lmax; l0 += 4) {
        double acc[4] = {0., 0., 0., 0.};
        for (int kk = 0; kk < KCH; kk++) {
            int k = blockIdx.z * KCH + kk;   // k = 0 .. nz+1
            if (!valid || k > g.nz + 1) continue;
            long q = gidx(g, i0, j0, k);
            T H = T(0.);
            if (l0 == 0) {
                H = heav_l<T>(ls, win, 0, q, i0, j0, k, eps, pi);
                for (int l = 1; l < lmax; l++) H = fmin(H, heav_l<T>(ls, win, l, q, i0, j0, k, eps, pi));
                miu[q] = miu1 * H + miu2 * (T(1.) - H);
                rho[q] = rho1 * H + rho2 * (T(1.) - H);
                if (jiemian) jiemian[q] = H;
            }
            if (k >= 1 && k <= g.nz) {
                for (int t = 0; t < 4 && l0 + t < lmax; t++)
                    acc[t] += (double)((l0 == 0 && t == 0 && lmax == 1) ? H
                                                                        : heav_l<T>(ls, win, l0 + t, q, i0, j0, k, eps, pi));
            }
        }
        for (int t = 0; t < 4 && l0 + t < lmax; t++) {
            double r = block_reduce(acc[t], OpSum(), smem);
            if (threadIdx.x == 0 && threadIdx.y == 0) part[(long)(l0 + t) * nblocks + block_linear()] = r;
        }
    }
}

// The same for one level set on the whole grid (nx a multiple of 256): four
// consecutive cells of a row per thread, moved as 16-byte vectors.  Per cell the
// operations are those of k_materials; the partial sums add the four cells of a layer
// before the next layer.
template <class T>
__global__ void __launch_bounds__(BX * BY)
k_materials1(Grid g, const T* __restrict__ phi, T eps, T pi, T miu1, T miu2, T rho1, T rho2, T* __restrict__ miu,
             T* __restrict__ rho, double* __restrict__ part) {
    __shared__ double smem[16];
    const int i0 = 4 * (blockIdx.x * BX + threadIdx.x), j0 = blockIdx.y * BY + threadIdx.y;
    const bool valid = j0 < g.ny;
    double acc = 0.;
    for (int kk = 0; kk < KCH; kk++) {
        const int k = blockIdx.z * KCH + kk;   // k = 0 .. nz+1
        if (!valid || k > g.nz + 1) continue;
        const long q = gidx(g, i0, j0, k);
        T p[4], mu[4], rh[4];
        ld4(phi + q, p);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const T H = heaviside_trig<T>(p[t], eps, pi);
            mu[t] = miu1 * H + miu2 * (T(1.) - H);
            rh[t] = rho1 * H + rho2 * (T(1.) - H);
            if (k >= 1 && k <= g.nz) acc += (double)H;
        }
        st4(miu + q, mu);
        st4(rho + q, rh);
    }
    const double r = block_reduce(acc, OpSum(), smem);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[block_linear()] = r;
}

// The same with level-set windows.  The block first lists (in LDS) the windows its
// cells can lie in; H is the minimum over those, and only their Heaviside sums are
// reduced.  Every other level set is 1 on all the block's cells: its partial sum is
// the block's interior cell count, the value the reduction of those 1s gives exactly.
// A block that meets more than MAT_LIST windows takes every level set.
constexpr int MAT_LIST = 128;
template <class T>
__global__ void k_materials_win(Grid g, int lmax, const T* const* __restrict__ ls, const int* __restrict__ win,
                                T eps, T pi, T miu1, T miu2, T rho1, T rho2, T* __restrict__ miu, T* __restrict__ rho,
                                double* __restrict__ part, int nblocks) {
    __shared__ double smem[16];
    __shared__ int list[MAT_LIST];
    __shared__ int nlist;
    const int tid = threadIdx.x + BX * threadIdx.y;
    const int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    const bool valid = i0 < g.nx && j0 < g.ny;
    const int xlo = blockIdx.x * BX, ylo = blockIdx.y * BY, klo = blockIdx.z * KCH;
    const int xhi = min(xlo + BX, g.nx) - 1, yhi = min(ylo + BY, g.ny) - 1, khi = min(klo + KCH - 1, g.nz + 1);
    if (tid == 0) nlist = 0;
    __syncthreads();
    for (int l = tid; l < lmax; l += BX * BY) {
        const int* w = win + 6 * l;
        if (w[0] <= xhi && w[0] + w[3] - 1 >= xlo && w[1] <= yhi && w[1] + w[4] - 1 >= ylo && w[2] - 2 <= khi &&
            w[2] + w[5] + 3 >= klo) {
            const int s = atomicAdd(&nlist, 1);
            if (s < MAT_LIST) list[s] = l;
        }
    }
    __syncthreads();
    const bool all = nlist > MAT_LIST;
    const int nl = all ? lmax : nlist;
    auto lid = [&](int e) { return all ? e : list[e]; };
    double cnt = 0.;
    for (int kk = 0; kk < KCH; kk++) {
        const int k = klo + kk;
        if (!valid || k > g.nz + 1) continue;
        const long q = gidx(g, i0, j0, k);
        T H = T(1.);
        for (int e = 0; e < nl; e++) H = fmin(H, heav_l<T>(ls, win, lid(e), q, i0, j0, k, eps, pi));
        miu[q] = miu1 * H + miu2 * (T(1.) - H);
        rho[q] = rho1 * H + rho2 * (T(1.) - H);
        if (k >= 1 && k <= g.nz) cnt += 1.;
    }
    __shared__ double count_s;
    const double c0 = block_reduce(cnt, OpSum(), smem);   // valid in thread 0
    if (tid == 0) count_s = c0;
    __syncthreads();
    const double count = count_s;
    for (int l = tid; l < lmax; l += BX * BY) part[(long)l * nblocks + block_linear()] = count;
    __syncthreads();   // the listed level sets' partials overwrite the counts below
    for (int e = 0; e < nl; e++) {
        const int l = lid(e);
        double acc = 0.;
        for (int kk = 0; kk < KCH; kk++) {
            const int k = klo + kk;
            if (!valid || k < 1 || k > g.nz) continue;
            acc += (double)heav_l<T>(ls, win, l, gidx(g, i0, j0, k), i0, j0, k, eps, pi);
        }
        const double r = block_reduce(acc, OpSum(), smem);
        if (tid == 0) part[(long)l * nblocks + block_linear()] = r;
    }
}

// volume(l) = 1 - sum/N (misc.f90:134), rho_avr (misc.f90:147)
__global__ void k_volume_final(int lmax, double ncell, double rho1, double rho2, const double* __restrict__ sums,
                               double* dscal) {
    if (threadIdx.x != 0) return;
    double sv = 0.;
    for (int l = 0; l < lmax; l++) {
        double v = 1. - sums[l] / ncell;
        dscal[DS_COUNT_FIXED + l] = v;
        sv = sv + v;
    }
    dscal[DS_RHO_AVR] = rho1 * (1. - sv) + rho2 * sv;
}

int launch_materials(Ctx& c) {
    ProfScope ps(c, "k_materials");
    dim3 blk(BX, BY, 1);
    dim3 grd = red_grid(c.g, 0, c.g.nz + 1);
    // one level set on the whole grid: the vector kernel, a quarter of the x blocks
    const bool vec = c.lmax == 1 && c.win.empty() && c.g.nx % (4 * BX) == 0;
    if (vec) grd.x = c.g.nx / (4 * BX);
    int nb = red_nblocks(grd);
    if ((long)nb * c.lmax > c.npart_cap) { set_error("reduction scratch too small"); return ICLS_ERR_STATE; }
    const icls_params& P = c.P;
    if (!c.win.empty()) {
        if (c.fp32)
            hipLaunchKernelGGL(k_materials_win<float>, grd, blk, 0, c.stream, c.g, c.lmax, (const float* const*)c.dls,
                               (const int*)c.dwin, (float)(1.5 * c.g.dz), (float)c.pi, (float)P.miu1, (float)P.miu2,
                               (float)P.rho1, (float)P.rho2, as<float>(c.miu), as<float>(c.rho), c.dpart, nb);
        else
            hipLaunchKernelGGL(k_materials_win<double>, grd, blk, 0, c.stream, c.g, c.lmax, (const double* const*)c.dls,
                               (const int*)c.dwin, 1.5 * c.g.dz, c.pi, P.miu1, P.miu2, P.rho1, P.rho2, c.miu, c.rho,
                               c.dpart, nb);
    } else if (vec) {
        if (c.fp32)
            hipLaunchKernelGGL(k_materials1<float>, grd, blk, 0, c.stream, c.g, as<float>(c.lvset[0]),
                               (float)(1.5 * c.g.dz), (float)c.pi, (float)P.miu1, (float)P.miu2, (float)P.rho1,
                               (float)P.rho2, as<float>(c.miu), as<float>(c.rho), c.dpart);
        else
            hipLaunchKernelGGL(k_materials1<double>, grd, blk, 0, c.stream, c.g, (const double*)c.lvset[0],
                               1.5 * c.g.dz, c.pi, P.miu1, P.miu2, P.rho1, P.rho2, c.miu, c.rho, c.dpart);
    } else if (c.fp32)
        hipLaunchKernelGGL(k_materials<float>, grd, blk, 0, c.stream, c.g, c.lmax, (const float* const*)c.dls,
                           (const int*)(c.win.empty() ? nullptr : c.dwin), (float)(1.5 * c.g.dz), (float)c.pi, (float)P.miu1, (float)P.miu2, (float)P.rho1,
                           (float)P.rho2, as<float>(c.miu), as<float>(c.rho), (float*)nullptr, c.dpart, nb);
    else
        hipLaunchKernelGGL(k_materials<double>, grd, blk, 0, c.stream, c.g, c.lmax, (const double* const*)c.dls,
                           (const int*)(c.win.empty() ? nullptr : c.dwin), 1.5 * c.g.dz, c.pi, P.miu1, P.miu2, P.rho1, P.rho2, c.miu, c.rho, (double*)nullptr, c.dpart,
                           nb);
    HIP_OK(hipGetLastError());
    double* sums = c.dscal + DS_COUNT_FIXED + 2 * c.lmax;   // lmax scratch sums after volume/init_volume
    hipLaunchKernelGGL(k_reduce_final_rows<OpSum>, dim3(c.lmax), dim3(RED_FINAL_THREADS), 0, c.stream,
                       (const double*)c.dpart, nb, 0., sums);
    double ncell = 1. * c.g.nx * c.g.ny * c.g.nz;
    hipLaunchKernelGGL(k_volume_final, dim3(1), dim3(64), 0, c.stream, c.lmax, ncell, P.rho1, P.rho2,
                       (const double*)sums, c.dscal);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// jiemian = min_l H_l over the C-halo range (misc.f90:115-120), on request only,
// written as double
template <class T>
__global__ void k_jiemian(Grid g, int lmax, const T* const* __restrict__ ls, T eps, T pi, double* __restrict__ out) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i0 >= g.nx || j0 >= g.ny) return;
    long q = gidx(g, i0, j0, k);
    T H = heaviside_trig<T>(ls[0][q], eps, pi);
    for (int l = 1; l < lmax; l++) H = fmin(H, heaviside_trig<T>(ls[l][q], eps, pi));
    out[q] = (double)H;
}

int launch_jiemian(Ctx& c, double* out) {
    dim3 blk(BX, BY, 1);
    dim3 grd = grid3(c.g, c.g.nz + 2, blk);
    if (c.fp32)
        hipLaunchKernelGGL(k_jiemian<float>, grd, blk, 0, c.stream, c.g, c.lmax, (const float* const*)c.dls,
                           (float)(1.5 * c.g.dz), (float)c.pi, out);
    else
        hipLaunchKernelGGL(k_jiemian<double>, grd, blk, 0, c.stream, c.g, c.lmax, (const double* const*)c.dls,
                           1.5 * c.g.dz, c.pi, out);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- interior sums
template <class T>
__global__ void k_sum_interior(Grid g, const T* __restrict__ f, double* __restrict__ part) {
    __shared__ double smem[16];
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    double acc = 0.;
    if (i0 < g.nx && j0 < g.ny)
        for (int kk = 0; kk < KCH; kk++) {
            int k = 1 + blockIdx.z * KCH + kk;
            if (k <= g.nz) acc += (double)f[gidx(g, i0, j0, k)];
        }
    double r = block_reduce(acc, OpSum(), smem);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[block_linear()] = r;
}

// dscal[slot] = sum over the interior of the field-class array f
int launch_sum_interior(Ctx& c, const double* f, int slot) {
    ProfScope ps(c, "k_sum_interior");
    dim3 blk(BX, BY, 1);
    dim3 grd = red_grid(c.g, 1, c.g.nz);
    int nb = red_nblocks(grd);
    if (c.fp32) hipLaunchKernelGGL(k_sum_interior<float>, grd, blk, 0, c.stream, c.g, as<float>(f), c.dpart);
    else hipLaunchKernelGGL(k_sum_interior<double>, grd, blk, 0, c.stream, c.g, f, c.dpart);
    hipLaunchKernelGGL(k_reduce_final<OpSum>, dim3(1), dim3(RED_FINAL_THREADS), 0, c.stream, (const double*)c.dpart, nb, 0.,
                       c.dscal + slot);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- ipot
// interface.f90:225-277: for the ordered pair (m, n), n != m, the minimum over every
// stored cell of phi_m + phi_n where -dz < phi_m < dz and phi_n < NarrowBand_2, 1e3
// elsewhere.  The x/y halos of the reference's arrays are periodic copies of stored
// cells, so the minimum over the stored cells (all z ghost layers included) is the
// reference's.  One block partial per partner n.
template <class T>
__global__ void k_ipot(Grid g, int lmax, int m, const T* const* __restrict__ ls, T dz, T cutoff,
                       double* __restrict__ part, int nblocks) {
    __shared__ double smem[16];
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    bool valid = i0 < g.nx && j0 < g.ny;
    for (int n = 0; n < lmax; n++) {
        if (n == m) continue;
        double d = 1e3;
        for (int kk = 0; kk < KCH; kk++) {
            int k = blockIdx.z * KCH + kk - 2;   // k = -2 .. nz+3
            if (!valid || k > g.nz + 3) continue;
            long q = gidx(g, i0, j0, k);
            T pm = ls[m][q], pn = ls[n][q];
            if (pm > -T(1.) * dz && pm < T(1.) * dz && pn < cutoff) d = fmin(d, (double)(pm + pn));
        }
        double r = block_reduce(d, OpMin(), smem);
        if (threadIdx.x == 0 && threadIdx.y == 0) part[(long)n * nblocks + block_linear()] = r;
    }
}

// dtemp[m + lmax n] for every ordered pair (device array of lmax * lmax doubles)
int launch_ipot(Ctx& c, double* dtemp) {
    ProfScope ps(c, "k_ipot");
    dim3 blk(BX, BY, 1);
    dim3 grd = red_grid(c.g, -2, c.g.nz + 3);
    int nb = red_nblocks(grd);
    if ((long)nb * c.lmax > c.npart_cap) { set_error("reduction scratch too small"); return ICLS_ERR_STATE; }
    for (int m = 0; m < c.lmax; m++) {
        if (c.fp32)
            hipLaunchKernelGGL(k_ipot<float>, grd, blk, 0, c.stream, c.g, c.lmax, m, (const float* const*)c.dls,
                               (float)c.g.dz, (float)c.NB2, c.dpart, nb);
        else
            hipLaunchKernelGGL(k_ipot<double>, grd, blk, 0, c.stream, c.g, c.lmax, m, (const double* const*)c.dls,
                               c.g.dz, c.NB2, c.dpart, nb);
        for (int n = 0; n < c.lmax; n++) {
            if (n == m) continue;
            hipLaunchKernelGGL(k_reduce_final<OpMin>, dim3(1), dim3(RED_FINAL_THREADS), 0, c.stream, (const double*)(c.dpart + (long)n * nb),
                               nb, 1e3, dtemp + m + (long)c.lmax * n);
        }
    }
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- mass correction
// ls_mass_correction trigger (levelset.f90:384-395): maxchk_glb = 0, then inflate
// when mod(istep,inflate_step) == 0 and |(V - V0)/V0| > trig (1e-5 in the reference).
__global__ void k_infl_flag(int l, int lmax, int due, double trig, double* dscal) {
    if (threadIdx.x != 0) return;
    dscal[DS_MAXCHK] = 0.;
    double V = dscal[DS_COUNT_FIXED + l], V0 = dscal[DS_COUNT_FIXED + lmax + l];
    double massloss = (V - V0) / V0;
    dscal[DS_INFL_FLAG] = (due && fabs(massloss) > trig) ? 1. : 0.;
}

// partial sums of delta*kappa*|grad phi| (levelset.f90:499), of the Heaviside
// (levelset.f90:505) and of delta^2*kappa*|grad phi|^2 (the normalisation of
// mass_correction = 2) over the interior; dheav over the C-halo range
template <class T>
__global__ void k_infl_sums(Grid g, const double* __restrict__ flag, const T* __restrict__ phi,
                            const T* __restrict__ kappa, const T* __restrict__ nm, T s1, T pi, T* __restrict__ dheav,
                            double* __restrict__ part, int nblocks) {
    if (flag[0] == 0.) return;
    __shared__ double smem[16];
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    bool valid = i0 < g.nx && j0 < g.ny;
    double sa = 0., sv = 0., s2 = 0.;
    for (int kk = 0; kk < KCH; kk++) {
        int k = blockIdx.z * KCH + kk;
        if (!valid || k > g.nz + 1) continue;
        long q = gidx(g, i0, j0, k);
        T h = heaviside_trig<T>(phi[q], s1, pi);
        dheav[q] = h;
        if (k >= 1 && k <= g.nz) {
            T del = dirac_trig<T>(phi[q], s1, pi);
            sa += (double)(del * kappa[q] * nm[q]);
            s2 += (double)(del * del * kappa[q] * (nm[q] * nm[q]));
            sv += (double)h;
        }
    }
    double ra = block_reduce(sa, OpSum(), smem);
    double rv = block_reduce(sv, OpSum(), smem);
    double r2 = block_reduce(s2, OpSum(), smem);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        part[block_linear()] = ra;
        part[nblocks + block_linear()] = rv;
        part[2 * nblocks + block_linear()] = r2;
    }
}

// levelset.f90:499-521 after the global sums.  mass_correction = 2 divides the loss
// rate by sum(delta^2 kappa |grad phi|^2) dV instead of the area: the correction
// velocity -temp kappa grad H moves phi by temp kappa delta |grad phi|^2 per unit time,
// which changes the H-measured volume at temp times that sum, so one correction
// restores the loss.  The reference's normalisation (mass_correction = 1) restores
// about int(delta^2) dphi = 3/(4 eps) = 1/(2 dz) times the loss.
// hoff: cells of the grid outside the swept window (Heaviside 1 there), 0 for a
// whole-grid sweep.  The three sums of k_infl_sums are each folded as k_reduce_final
// folds one array (the same order, the same results), then one thread forms the
// scalars.
__global__ void __launch_bounds__(RED_FINAL_THREADS)
k_infl_final(const double* __restrict__ part, int nb, int l, int lmax, int norm, double ncell, double hoff, double dx,
             double dy, double dz, double lxyz, double inflate_dt, double s1, double* __restrict__ sums,
             double* dscal) {
    if (dscal[DS_INFL_FLAG] == 0.) return;
    __shared__ double smem[16];
    double r[3];
#pragma unroll
    for (int t = 0; t < 3; t++) r[t] = block_reduce(fold_partials<OpSum>(part + (long)t * nb, nb, 0.), OpSum(), smem);
    if (threadIdx.x != 0) return;
    for (int t = 0; t < 3; t++) sums[t] = r[t];
    double area = (norm ? r[2] : r[0]) * dx * dy * dz;
    double vol = 1. - (r[1] + hoff) / ncell;
    dscal[DS_COUNT_FIXED + l] = vol;
    double loss_rate = lxyz * (vol - dscal[DS_COUNT_FIXED + lmax + l]) / inflate_dt;
    double temp = loss_rate / area;
    dscal[DS_AREA] = area;
    dscal[DS_INFL_TEMP] = temp;
    dscal[DS_MAXCHK] = temp / s1;
}

// u_l = -temp*(dheav(i+1)-dheav(i))/dx etc. (levelset.f90:527-535).  boundp(dheav) is
// read in place: dheav(k1) = dheav(kmax) at the top layer, whose w_l is then zero, the
// value bounduvw gives it (bound.f90:26-47).
template <class T>
__global__ void k_infl_face_vel(Grid g, const double* __restrict__ dscal, const T* __restrict__ dheav,
                                T* __restrict__ ul, T* __restrict__ vl, T* __restrict__ wl) {
    if (dscal[DS_INFL_FLAG] == 0.) return;
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    const T temp = (T)dscal[DS_INFL_TEMP];
    long q = gidx(g, i0, j0, k);
    T h = dheav[q];
    ul[q] = -temp * (dheav[gidx(g, wrap(i0 + 1, g.nx), j0, k)] - h) / T(g.dx);
    vl[q] = -temp * (dheav[gidx(g, i0, wrap(j0 + 1, g.ny), k)] - h) / T(g.dy);
    wl[q] = k < g.nz ? -temp * (dheav[gidx(g, i0, j0, k + 1)] - h) / T(g.dz) : T(0.);
}

// u_lc = (u_l(i-1)+u_l(i))/2*kappa (levelset.f90:541-549); below the first layer w_l is
// the wall value zero that bounduvw gives it
template <class T>
__global__ void k_infl_center_vel(Grid g, const double* __restrict__ dscal, const T* __restrict__ kappa,
                                  const T* __restrict__ ul, const T* __restrict__ vl, const T* __restrict__ wl,
                                  T* __restrict__ uc, T* __restrict__ vc, T* __restrict__ wc) {
    if (dscal[DS_INFL_FLAG] == 0.) return;
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    long q = gidx(g, i0, j0, k);
    T kap = kappa[q];
    uc[q] = (ul[gidx(g, wrap(i0 - 1, g.nx), j0, k)] + ul[q]) / T(2.) * kap;
    vc[q] = (vl[gidx(g, i0, wrap(j0 - 1, g.ny), k)] + vl[q]) / T(2.) * kap;
    wc[q] = ((k > 1 ? wl[gidx(g, i0, j0, k - 1)] : T(0.)) + wl[q]) / T(2.) * kap;
}

int launch_inflation_prepare(Ctx& c, int l, int istep) {
    ProfScope ps(c, "k_infl_flag");
    int due = (c.P.inflate_step > 0 && istep % c.P.inflate_step == 0) ? 1 : 0;
    hipLaunchKernelGGL(k_infl_flag, dim3(1), dim3(64), 0, c.stream, l, c.lmax, due, mass_trigger(c.P), c.dscal);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// Everything of inflation (levelset.f90:407-549) up to the cell-centre correction
// velocities; the caller has filled c.curv2 (kappa) and c.normal_m from the current
// level set with get_curvature predicated on the same flag.
template <class T>
static int inflation_velocity(Ctx& c, int l) {
    const double* flag = c.dscal + DS_INFL_FLAG;
    const double s1 = 1.5 * c.g.dz;
    dim3 blk(BX, BY, 1);
    dim3 grd = red_grid(c.g, 0, c.g.nz + 1);
    int nb = red_nblocks(grd);
    if (3L * nb > c.npart_cap) { set_error("reduction scratch too small"); return ICLS_ERR_STATE; }
    hipLaunchKernelGGL(k_infl_sums<T>, grd, blk, 0, c.stream, c.g, flag, as<T>(c.lvset[l]), as<T>(c.curv2),
                       as<T>(c.normal_m), (T)s1, (T)c.pi, as<T>(c.dheav), c.dpart, nb);
    double* sums = c.dscal + DS_COUNT_FIXED + 2 * c.lmax;
    double ncell = 1. * c.ncell_all;
    double hoff = (double)(c.ncell_all - (long)c.g.nx * c.g.ny * c.g.nz);
    double lxyz = c.P.lx * c.P.ly * c.P.lz;
    double inflate_dt = c.P.inflate_step * c.dt;
    hipLaunchKernelGGL(k_infl_final, dim3(1), dim3(RED_FINAL_THREADS), 0, c.stream, (const double*)c.dpart, nb, l,
                       c.lmax, c.P.mass_correction == 2 ? 1 : 0, ncell, hoff, c.g.dx, c.g.dy, c.g.dz, lxyz,
                       inflate_dt, s1, sums, c.dscal);
    hipLaunchKernelGGL(k_infl_face_vel<T>, grid3(c.g, c.g.nz, blk), blk, 0, c.stream, c.g, (const double*)c.dscal,
                       as<T>(c.dheav), as<T>(c.u_l), as<T>(c.v_l), as<T>(c.w_l));
    HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(k_infl_center_vel<T>, grid3(c.g, c.g.nz, blk), blk, 0, c.stream, c.g, (const double*)c.dscal,
                       as<T>(c.curv2), as<T>(c.u_l), as<T>(c.v_l), as<T>(c.w_l), as<T>(c.ucc_l), as<T>(c.vcc_l),
                       as<T>(c.wcc_l));
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

int launch_inflation_velocity(Ctx& c, int l) {
    ProfScope ps(c, "k_infl");
    return c.fp32 ? inflation_velocity<float>(c, l) : inflation_velocity<double>(c, l);
}

}  // namespace icls
