// k_levelset.hip -- narrow-band level-set kernels: tube construction, SSP-RK3
// advection stages (HOUC5 / WENO5), WENO5 re-initialisation stages, z-wall
// extrapolation, initial sphere SDFs.
//
// Every kernel sweeps the whole padded field with one thread per cell; x and y wrap
// periodically, which replaces the 3-cell halo exchange updthalos_ls
// (bound.f90:760-797).  The narrow band is a byte mask over the interior instead of
// the reference's compacted index lists (levelset.f90:160-188); a dense masked sweep
// touches exactly the cells RK_nb updates (levelset.f90:353-361).
//
// T is the element type of the level-set arrays (double, or float in the mixed
// modes).  Every literal is written as T(...), so the double instantiation performs
// exactly the operations of the reference in its order.
#include "icls_internal.hpp"
#include "icls_reduce.hpp"
#include "icls_stencil.hpp"

namespace icls {

// Tube and re-init tiles go to the XCDs in groups of 8 (xcd_tile_dealt): the band, where
// their work is, lies in a few z chunks, which contiguous ranges would give to two or
// three XCDs.  At 256^3 (MI355X) groups of 2, 4, 8 and 16 ran the float re-init stage
// 2/3 in 92.7, 88.1, 87.8 and 88.0 us against 98.5 with contiguous ranges (FP64: 170
// against 229); the tube was unchanged.
constexpr int XCD_DEAL = 8;

static constexpr int BX = 64, BY = 4;

// ---------------------------------------------------------------- tube_construction
// Clamp phi to +-NarrowBand_2 over the whole array (levelset.f90:107-119) and mark
// the band: a cell is in the band unless all 13 points of its +-2 axis scans sit at
// |phi| == NarrowBand_2 (levelset.f90:125-156; a sum of non-positive terms is zero
// only when every term is).  A clamped value has |phi| < NarrowBand_2 exactly when
// the unclamped one has, so the scan tests the input directly.
// The boundsls that follows (bound.f90:404-411) is folded in: layers 0 and k1 take
// the linear extrapolation of the clamped layers next to them.
//
// The construction marches in z over a 64 x 4 tile: the input passes through a
// ring of eight LDS planes with a two-cell x/y halo (the five planes k-2 .. k+2 of
// step k, and free slots for the next planes, so a step needs one barrier), the
// 13-point scan reads LDS and every input element is fetched once per block.  Chunks of TZ layers;
// the first and last chunk also write the ghost layers.  TPD planes are loaded ahead
// (unconditional loads, clamped to the chunk, as in the re-initialisation kernel).
// Chunks of TZ_GRID layers on the whole grid, of TZ_WINDOW inside a level-set window.
constexpr int TZ_GRID = 32, TZ_WINDOW = 8, TH = 2, TTW = BX + 2 * TH, TTH = BY + 2 * TH, TTP = TTW * TTH;
constexpr int TNLD = (TTP + BX * BY - 1) / (BX * BY), TNS = 8;   // ring slots: 5 planes of a step and the next

template <class T, int TPD>
__device__ __forceinline__ void tube_zm_body(const Grid& g, int bid, int TZ, T NB2, const T* __restrict__ in,
                                             T* __restrict__ out, uint8_t* __restrict__ mask) {
    __shared__ T ring[TNS][TTP];
    const int ntx = (g.nx + BX - 1) / BX, nty = (g.ny + BY - 1) / BY, ntz = (g.nz + TZ - 1) / TZ;
    const int tile = xcd_tile_dealt(bid, ntx * nty * ntz, XCD_DEAL);
    const int bx = tile % ntx, by = (tile / ntx) % nty, bz = tile / (ntx * nty);
    const int x0 = bx * BX, y0 = by * BY, kb = 1 + bz * TZ;
    const int ke = kb + TZ - 1 < g.nz ? kb + TZ - 1 : g.nz;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + BX * ty;
    const int xm = g.nx - 1;   // nx is a power of two; ny need not be (a window's)
    // a thread past the end of x or y takes the wrapped cell (same values, same stores)
    const int i = (x0 + tx) & xm, j = wrap(y0 + ty, g.ny);
    long poff[TNLD];
#pragma unroll
    for (int r = 0; r < TNLD; r++) {
        int e = tid + r * BX * BY;
        if (e >= TTP) e = TTP - 1;
        const int lx = e % TTW - TH, ly = e / TTW - TH;
        poff[r] = (long)((x0 + lx) & xm) + g.sy * (long)wrap(y0 + ly, g.ny);
    }
    T nxt[TPD][TNLD];
    auto fetch = [&](int k, T* dst) {
        const long base = g.sz * (long)(k + 2);
#pragma unroll
        for (int r = 0; r < TNLD; r++) dst[r] = in[base + poff[r]];
    };
    auto slot_of = [&](int k) { return (k - kb + 2 * TNS) & (TNS - 1); };
    auto stash_plane = [&](int k, const T* src) {
        T* dst = ring[slot_of(k)];
#pragma unroll
        for (int r = 0; r < TNLD; r++) {
            const int e = tid + r * BX * BY;
            if (e < TTP) dst[e] = src[r];
        }
    };
    auto clampv = [&](T v) { return v > NB2 ? NB2 : (v < -NB2 ? -NB2 : v); };
    auto col = [&](int k) { return in[gidx(g, i, j, k)]; };
    // ghost layers below the first chunk: clamped, layer 0 extrapolated
    if (bz == 0) {
        out[gidx(g, i, j, -2)] = clampv(col(-2));
        out[gidx(g, i, j, -1)] = clampv(col(-1));
        out[gidx(g, i, j, 0)] = T(2.) * clampv(col(1)) - clampv(col(2));
    }
    {
        // the four planes before the first step: all loads in flight before the first store
        T pro[2 * TH][TNLD];
#pragma unroll
        for (int t = 0; t < 2 * TH; t++) fetch(kb - TH + t, pro[t]);
#pragma unroll
        for (int t = 0; t < 2 * TH; t++) stash_plane(kb - TH + t, pro[t]);
    }
#pragma unroll
    for (int u = 0; u < TPD; u++) fetch(min(kb + u, ke) + TH, nxt[u]);
    const int c0 = (ty + TH) * TTW + tx + TH;
    for (int k0 = kb; k0 <= ke; k0 += TPD) {
#pragma unroll
        for (int u = 0; u < TPD; u++) {
            const int k = k0 + u;
            stash_plane(k + TH, nxt[u]);
            __syncthreads();
            fetch(min(k + TPD, ke) + TH, nxt[u]);
            const T* P = ring[slot_of(k)];
            const T v = P[c0];
            bool m = fabs(v) < NB2;
#pragma unroll
            for (int d = 1; d <= TH; d++) {
                m = m || fabs(P[c0 - d]) < NB2 || fabs(P[c0 + d]) < NB2;
                m = m || fabs(P[c0 - d * TTW]) < NB2 || fabs(P[c0 + d * TTW]) < NB2;
                m = m || fabs(ring[slot_of(k - d)][c0]) < NB2 || fabs(ring[slot_of(k + d)][c0]) < NB2;
            }
            const long q = gidx(g, i, j, k);
            out[q] = clampv(v);
            mask[q] = m ? 1 : 0;
            // no barrier here: the next step stores plane k + 3 into the slot of plane
            // k - 5, which no thread reads after the barrier above
        }
    }
    if (bz == ntz - 1) {
        out[gidx(g, i, j, g.nz + 1)] = T(2.) * clampv(col(g.nz)) - clampv(col(g.nz - 1));
        out[gidx(g, i, j, g.nz + 2)] = clampv(col(g.nz + 2));
        out[gidx(g, i, j, g.nz + 3)] = clampv(col(g.nz + 3));
    }
}

template <class T, int TPD>
__global__ void __launch_bounds__(BX * BY)
k_tube_zm(Grid g, int TZ, T NB2, const T* __restrict__ in, T* __restrict__ out, uint8_t* __restrict__ mask) {
    tube_zm_body<T, TPD>(g, blockIdx.x, TZ, NB2, in, out, mask);
}

// the same over the windows of a batch (g: the whole grid)
template <class T, int TPD>
__global__ void __launch_bounds__(BX * BY)
k_tube_zm_batch(Grid g, WinBatch wb, int TZ, T NB2, const T* in, T* out, uint8_t* mask) {
    const WinPick w = win_pick(g, wb, blockIdx.x);
    tube_zm_body<T, TPD>(w.g, w.bid, TZ, NB2, win_ptr(in, wb, w), win_ptr(out, wb, w), mask + w.off);
}

template <class T>
static void tube_dispatch(Ctx& c, const double* phi_in, double* phi_out) {
    dim3 blk(BX, BY, 1);
    const int tz = in_window(c) ? TZ_WINDOW : TZ_GRID;
    const int ntiles = ((c.g.nx + BX - 1) / BX) * ((c.g.ny + BY - 1) / BY) * ((c.g.nz + tz - 1) / tz);
    // groups of 2 steps when every chunk length is even
    if (tz % 2 == 0 && (c.g.nz % tz) % 2 == 0)
        hipLaunchKernelGGL((k_tube_zm<T, 2>), dim3(ntiles), blk, 0, c.stream, c.g, tz, (T)c.NB2, as<T>(phi_in),
                           as<T>(phi_out), c.mask);
    else
        hipLaunchKernelGGL((k_tube_zm<T, 1>), dim3(ntiles), blk, 0, c.stream, c.g, tz, (T)c.NB2, as<T>(phi_in),
                           as<T>(phi_out), c.mask);
}

int launch_tube(Ctx& c, const double* phi_in, double* phi_out) {
    ProfScope ps(c, "k_tube");
    if (c.fp32) tube_dispatch<float>(c, phi_in, phi_out);
    else tube_dispatch<double>(c, phi_in, phi_out);
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

template <class T>
__global__ void k_bc_ls(Grid g, T* __restrict__ phi, const T* __restrict__ res_ref, double* res,
                        const double* __restrict__ flag) {
    if (flag && flag[0] == 0.) return;
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    bool valid = i0 < g.nx && j0 < g.ny;
    double r = 0.;
    if (valid) {
        const int km = g.nz;
        T lo = T(2.) * phi[gidx(g, i0, j0, 1)] - phi[gidx(g, i0, j0, 2)];
        T hi = T(2.) * phi[gidx(g, i0, j0, km)] - phi[gidx(g, i0, j0, km - 1)];
        phi[gidx(g, i0, j0, 0)] = lo;
        phi[gidx(g, i0, j0, km + 1)] = hi;
        // res_ref: the two stashed wall layers of phi0, plane 0 then plane k1
        if (res_ref) r = (double)fmax(fabs(lo - res_ref[i0 + g.sy * j0]), fabs(hi - res_ref[g.sz + i0 + g.sy * j0]));
    }
    if (res_ref) wave_atomic_max(res, r);
}

static int bc_ls(Ctx& c, double* phi, const double* res_ref, double* res, const double* flag) {
    ProfScope ps(c, "k_bc_ls");
    dim3 blk(BX, BY, 1);
    dim3 grd = grid3(c.g, 1, blk);
    if (c.fp32)
        hipLaunchKernelGGL(k_bc_ls<float>, grd, blk, 0, c.stream, c.g, as<float>(phi), as<float>(res_ref), res, flag);
    else
        hipLaunchKernelGGL(k_bc_ls<double>, grd, blk, 0, c.stream, c.g, phi, res_ref, res, flag);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// boundmls: three extrapolated layers per wall (bound.f90:539-549), init only
template <class T>
__global__ void k_boundmls(Grid g, T* __restrict__ y) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    if (i0 >= g.nx || j0 >= g.ny) return;
    const int kmax = g.nz, k1 = g.nz + 1;
    auto Y = [&](int k) -> T& { return y[gidx(g, i0, j0, k)]; };
    Y(0) = T(2.) * Y(1) - Y(2);
    Y(-1) = T(2.) * Y(0) - Y(1);
    Y(-2) = T(2.) * Y(-1) - Y(0);
    Y(k1) = T(2.) * Y(kmax) - Y(kmax - 1);
    Y(k1 + 1) = T(2.) * Y(k1) - Y(kmax);
    Y(k1 + 2) = T(2.) * Y(k1 + 1) - Y(k1);
}

int launch_boundmls(Ctx& c, double* phi) {
    ProfScope ps(c, "k_boundmls");
    dim3 blk(BX, BY, 1);
    if (c.fp32) hipLaunchKernelGGL(k_boundmls<float>, grid3(c.g, 1, blk), blk, 0, c.stream, c.g, as<float>(phi));
    else hipLaunchKernelGGL(k_boundmls<double>, grid3(c.g, 1, blk), blk, 0, c.stream, c.g, phi);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- advection stage
// dfdt of HOUC5_nb (hyperPDE.f90:397-495) or WENO5_nb (hyperPDE.f90:503-800) at one
// band cell.  The z stencil drops to HOUC3 at k = 2, kmax-1 and to first order at
// k = 1, kmax in both schemes.
template <class T, int SCHEME, bool P2>
__device__ __forceinline__ T adv_rhs(const Grid& g, const T* __restrict__ phi, int i0, int j0, int k, T vu, T vv, T vw) {
    T f[7];
    const Sp<T, P2> sx{T(g.dx), T(g.dxi)}, sy{T(g.dy), T(g.dyi)}, sz{T(g.dz), T(g.dzi)};
#pragma unroll
    for (int d = -3; d <= 3; d++) f[d + 3] = phi[gidx(g, wrap(i0 + d, g.nx), j0, k)];
    T udfdx, vdfdy, wdfdz;
    if (SCHEME == ICLS_ADV_HOUC5) {
        udfdx = sx.div(-vu * houc5<T>(vu, f + 3));
    } else {
        T dfx = (vu >= T(0.)) ? weno_minus(f + 3, sx) : weno_plus(f + 3, sx);
        udfdx = -vu * dfx;
    }
#pragma unroll
    for (int d = -3; d <= 3; d++) f[d + 3] = phi[gidx(g, i0, wrap(j0 + d, g.ny), k)];
    if (SCHEME == ICLS_ADV_HOUC5) {
        vdfdy = sy.div(-vv * houc5<T>(vv, f + 3));
    } else {
        // y differences are divided by dx in the reference (hyperPDE.f90:622-626)
        T dfy = (vv >= T(0.)) ? weno_minus(f + 3, sx) : weno_plus(f + 3, sx);
        vdfdy = -vv * dfy;
    }
    const int kmax = g.nz;
    if (k > 2 && k < kmax - 1) {
#pragma unroll
        for (int d = -3; d <= 3; d++) f[d + 3] = phi[gidx(g, i0, j0, k + d)];
        if (SCHEME == ICLS_ADV_HOUC5) {
            wdfdz = sz.div(-vw * houc5<T>(vw, f + 3));
        } else {
            T dfz = (vw >= T(0.)) ? weno_minus(f + 3, sz) : weno_plus(f + 3, sz);
            wdfdz = -vw * dfz;
        }
    } else {
#pragma unroll
        for (int d = -2; d <= 2; d++) f[d + 3] = phi[gidx(g, i0, j0, k + d)];
        T dfz = (k == 2 || k == kmax - 1) ? houc3<T>(vw, f + 3) : upwind1<T>(vw, f + 3);
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
// The boundsls that follows every stage (bound.f90:404-411) is folded in: the
// threads of layers 0 and k1 evaluate the stage at the two layers next to the wall
// (the same operations the threads of those layers perform) and store the linear
// extrapolation.
template <class T, int SCHEME, int STAGE, bool P2, bool VELF>
__device__ __forceinline__ void adv_stage_body(const Grid& g, int bx, int by, int ks, T dt, const T* phi0,
                                               const T* __restrict__ phis, T* out, const uint8_t* __restrict__ mask,
                                               const T* __restrict__ vu, const T* __restrict__ vv,
                                               const T* __restrict__ vw) {
    int i0 = bx * BX + threadIdx.x, j0 = by * BY + threadIdx.y;
    if (i0 >= g.nx || j0 >= g.ny) return;
    int k = ks - 2;
    auto value = [&](int kk) -> T {
        const long qq = gidx(g, i0, j0, kk);
        T comb, factor;
        if (STAGE == 1) { comb = phi0[qq]; factor = T(1.); }
        else if (STAGE == 2) { comb = T(3.) / T(4.) * phi0[qq] + T(1.) / T(4.) * phis[qq]; factor = T(1.) / T(4.); }
        else { comb = T(1.) / T(3.) * phi0[qq] + T(2.) / T(3.) * phis[qq]; factor = T(2.) / T(3.); }
        T r = comb;
        if (kk >= 1 && kk <= g.nz && mask[qq]) {
            T uc, vc, wc;
            if (VELF) {
                uc = (vu[gidx(g, wrap(i0 - 1, g.nx), j0, kk)] + vu[qq]) / T(2.);
                vc = (vv[gidx(g, i0, wrap(j0 - 1, g.ny), kk)] + vv[qq]) / T(2.);
                wc = (vw[gidx(g, i0, j0, kk - 1)] + vw[qq]) / T(2.);
            } else {
                uc = vu[qq]; vc = vv[qq]; wc = vw[qq];
            }
            T dfdt = adv_rhs<T, SCHEME, P2>(g, phis, i0, j0, kk, uc, vc, wc);
            r = comb + factor * dt * dfdt;
        }
        return r;
    };
    T r;
    if (k == 0) r = T(2.) * value(1) - value(2);
    else if (k == g.nz + 1) r = T(2.) * value(g.nz) - value(g.nz - 1);
    else r = value(k);
    out[gidx(g, i0, j0, k)] = r;
}

// The same stage with four consecutive cells of a row per thread on the whole grid (nx
// a multiple of 256): phi0, phi_s, the mask and the result move as vectors; a band cell
// takes the stencil of adv_stage_body, with the same operations.
template <class T, int SCHEME, int STAGE, bool P2, bool VELF>
__global__ void __launch_bounds__(BX * BY)
k_adv_stage4(Grid g, T dt, const T* phi0, const T* __restrict__ phis, T* out, const uint8_t* __restrict__ mask,
             const T* __restrict__ vu, const T* __restrict__ vv, const T* __restrict__ vw,
             const double* __restrict__ flag) {
    if (flag && flag[0] == 0.) return;
    const int i0 = 4 * (blockIdx.x * BX + threadIdx.x), j0 = blockIdx.y * BY + threadIdx.y;
    if (j0 >= g.ny) return;
    const int k = (int)blockIdx.z - 2;
    T factor, comb[4], p0[4], ps[4] = {T(0.), T(0.), T(0.), T(0.)};
    auto comb_of = [&](T a, T b) -> T {
        if (STAGE == 1) return a;
        if (STAGE == 2) return T(3.) / T(4.) * a + T(1.) / T(4.) * b;
        return T(1.) / T(3.) * a + T(2.) / T(3.) * b;
    };
    if (STAGE == 1) factor = T(1.);
    else if (STAGE == 2) factor = T(1.) / T(4.);
    else factor = T(2.) / T(3.);
    auto dfdt_at = [&](int ii, int kk, long qq) -> T {
        T uc, vc, wc;
        if (VELF) {
            uc = (vu[gidx(g, wrap(ii - 1, g.nx), j0, kk)] + vu[qq]) / T(2.);
            vc = (vv[gidx(g, ii, wrap(j0 - 1, g.ny), kk)] + vv[qq]) / T(2.);
            wc = (vw[gidx(g, ii, j0, kk - 1)] + vw[qq]) / T(2.);
        } else {
            uc = vu[qq]; vc = vv[qq]; wc = vw[qq];
        }
        return adv_rhs<T, SCHEME, P2>(g, phis, ii, j0, kk, uc, vc, wc);
    };
    // the four cells of layer kk, as adv_stage_body's value()
    auto values = [&](int kk, T* r) {
        const long qq = gidx(g, i0, j0, kk);
        ld4(phi0 + qq, p0);
        if (STAGE > 1) ld4(phis + qq, ps);
        uchar4 m = {0, 0, 0, 0};
        if (kk >= 1 && kk <= g.nz) m = *reinterpret_cast<const uchar4*>(mask + qq);
        const unsigned char mm[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
        for (int t = 0; t < 4; t++) {
            comb[t] = comb_of(p0[t], ps[t]);
            r[t] = comb[t];
            if (mm[t]) r[t] = comb[t] + factor * dt * dfdt_at(i0 + t, kk, qq + t);
        }
    };
    T r[4];
    if (k == 0 || k == g.nz + 1) {
        T a[4], b[4];
        values(k == 0 ? 1 : g.nz, a);
        values(k == 0 ? 2 : g.nz - 1, b);
#pragma unroll
        for (int t = 0; t < 4; t++) r[t] = T(2.) * a[t] - b[t];
    } else {
        values(k, r);
    }
    st4(out + gidx(g, i0, j0, k), r);
}

template <class T, int SCHEME, int STAGE, bool P2, bool VELF>
__global__ void k_adv_stage(Grid g, T dt, const T* phi0, const T* __restrict__ phis, T* out,
                            const uint8_t* __restrict__ mask, const T* __restrict__ vu, const T* __restrict__ vv,
                            const T* __restrict__ vw, const double* __restrict__ flag) {
    if (flag && flag[0] == 0.) return;
    adv_stage_body<T, SCHEME, STAGE, P2, VELF>(g, blockIdx.x, blockIdx.y, blockIdx.z, dt, phi0, phis, out, mask, vu,
                                               vv, vw);
}

// the same over the windows of a batch (g: the whole grid); the blocks of a window
// are its (x, y, layer) blocks of the one-window launch, x fastest
template <class T, int SCHEME, int STAGE, bool P2, bool VELF>
__global__ void __launch_bounds__(BX * BY)
k_adv_stage_batch(Grid g, WinBatch wb, T dt, const T* phi0, const T* phis, T* out, const uint8_t* mask,
                  const T* vu, const T* vv, const T* vw) {
    const WinPick w = win_pick(g, wb, blockIdx.x);
    const int nbx = (w.g.nx + BX - 1) / BX, nby = (w.g.ny + BY - 1) / BY;
    adv_stage_body<T, SCHEME, STAGE, P2, VELF>(w.g, w.bid % nbx, (w.bid / nbx) % nby, w.bid / (nbx * nby), dt,
                                               win_ptr(phi0, wb, w), win_ptr(phis, wb, w), win_ptr(out, wb, w),
                                               mask + w.off, vu + w.off, vv + w.off, vw + w.off);
}

template <class T>
static void adv_dispatch(Ctx& c, int scheme, int stage, const double* phi0, const double* phis, double* out,
                         const double* vu, const double* vv, const double* vw, int velf, const double* flag) {
    dim3 blk(BX, BY, 1);
    dim3 grd = grid3(c.g, c.g.nz + 6, blk);
    // four cells per thread on the whole grid (the arrays of a level-set window start
    // where the window does, not on a vector boundary)
    auto al = [](const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; };
    const bool vec = !in_window(c) && c.g.nx % (4 * BX) == 0 && c.g.sy % 4 == 0 && al(phi0) && al(phis) &&
                     al(out) && reinterpret_cast<uintptr_t>(c.mask) % 4 == 0;
    if (vec) grd.x = c.g.nx / (4 * BX);
#define ADV1(S, ST, P, V)                                                                                          \
    do {                                                                                                           \
        if (vec)                                                                                                   \
            hipLaunchKernelGGL((k_adv_stage4<T, S, ST, P, V>), grd, blk, 0, c.stream, c.g, (T)c.dt, as<T>(phi0),  \
                               as<T>(phis), as<T>(out), (const uint8_t*)c.mask, as<T>(vu), as<T>(vv), as<T>(vw),  \
                               flag);                                                                              \
        else                                                                                                       \
            hipLaunchKernelGGL((k_adv_stage<T, S, ST, P, V>), grd, blk, 0, c.stream, c.g, (T)c.dt, as<T>(phi0),   \
                               as<T>(phis), as<T>(out), (const uint8_t*)c.mask, as<T>(vu), as<T>(vv), as<T>(vw),  \
                               flag);                                                                              \
    } while (0)
#define ADV(S, ST)                                                             \
    if (scheme == S && stage == ST) {                                          \
        if (c.g.pow2) {                                                        \
            if (velf) ADV1(S, ST, true, true); else ADV1(S, ST, true, false);   \
        } else {                                                               \
            if (velf) ADV1(S, ST, false, true); else ADV1(S, ST, false, false); \
        }                                                                      \
    }
    ADV(ICLS_ADV_HOUC5, 1) ADV(ICLS_ADV_HOUC5, 2) ADV(ICLS_ADV_HOUC5, 3)
    ADV(ICLS_ADV_WENO5, 1) ADV(ICLS_ADV_WENO5, 2) ADV(ICLS_ADV_WENO5, 3)
#undef ADV
#undef ADV1
}

int launch_adv_stage(Ctx& c, int scheme, int stage, const double* phi0, const double* phis, double* out,
                     const double* vu, const double* vv, const double* vw, int velf, const double* flag) {
    ProfScope ps(c, stage == 1 ? "k_adv_stage1" : "k_adv_stage23");
    if (c.fp32) adv_dispatch<float>(c, scheme, stage, phi0, phis, out, vu, vv, vw, velf, flag);
    else adv_dispatch<double>(c, scheme, stage, phi0, phis, out, vu, vv, vw, velf, flag);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- re-initialisation
// One RK3 stage of ls_reinit (levelset.f90:798-810) with the smoothed sign of
// levelset.f90:668-693 recomputed from phi0 on the fly:
//   stage 1: out = phi0 + dtau*RHS(phi0)
//   stage 2: out = 3/4 phi0 + 1/4 (phi_n1 + dtau*RHS(phi_n1))
//   stage 3: out = 1/3 phi0 + 2/3 (phi_n2 + dtau*RHS(phi_n2))
// RHS is WENO5_reinit (hyperPDE.f90:2662-2736, 2932-3197): six one-sided WENO5
// derivatives and the Godunov-type selection, evaluated where |phi| < NarrowBand_2 +
// dz and zero elsewhere and in the ghost layers.  Along an axis whose seven values are
// equal (the clamped far field) both WENO5 derivatives are exactly +0 (every
// difference is +0, every smoothness indicator 0, the weighted sum of zeros +0), so
// they are set without evaluating the weights.
// The boundsls after each stage (bound.f90:404-411) is folded in when the last chunk
// holds two layers or more (WALL): the block that computes layers 1 and 2 (kmax-1 and
// kmax) stores the linear extrapolation into layer 0 (k1).  Stage 3 also folds |out -
// phi0| of every layer into the residual (levelset.f90:831), the two wall layers with
// their extrapolated values; without WALL it stashes phi0 of those layers for
// k_bc_ls.  Stage 1 resets the residual.  Stage 3 may run in place (out == phi0).
//
// The kernel marches in z over a 64 x 4 tile of columns.  phi_s passes through
// a ring of eight LDS planes of the tile with a three-cell x/y halo: the seven planes
// k-3 .. k+3 of step k and a free slot, which takes plane k+4 at the start of step
// k+1 while slower wavefronts may still read k-3, so a step needs one barrier.  The
// 19-point WENO stencils read LDS and each phi_s element is fetched once per block;
// the next planes are loaded into registers while the current one is computed.
// Uniform windows: when a plane is stored, the block also records whether all its
// values are one value.  Where the seven planes of a step are one value v (the
// clamped far field, most of the domain) every stencil is flat, all six WENO5
// derivatives are +0, and RHS = s exactly wherever |v| < NarrowBand_2 + dz; the step
// then skips the stencils.
// Interior layers are split into chunks of rz (zchunk_for); the first and the last
// chunk also write the three ghost layers below and above.  Tiles are taken in XCD
// order.
#ifndef ICLS_REINIT_PD
#define ICLS_REINIT_PD 4
#endif
#ifndef ICLS_REINIT_RR
#define ICLS_REINIT_RR 1
#endif
constexpr int RH = 3;                                    // stencil radius
constexpr int RR = ICLS_REINIT_RR;                       // rows per thread
constexpr int RTY = BY * RR;                             // tile rows
constexpr int RTW = BX + 2 * RH, RTH = RTY + 2 * RH, RTP = RTW * RTH;
constexpr int RNLD = (RTP + BX * BY - 1) / (BX * BY);
constexpr int RNS = 8;                                    // ring slots: the 7 planes of a step and the next
constexpr int RPD_MAX = ICLS_REINIT_PD;                   // planes in flight ahead of the ring

template <class T, int STAGE, bool P2>
__device__ __forceinline__ T reinit_combine(T p0, T ps, T rhs, T dtau) {
    if (STAGE == 1) return p0 + dtau * rhs;
    if (STAGE == 2) return T(3.) / T(4.) * p0 + T(1.) / T(4.) * (ps + dtau * rhs);
    return T(1.) / T(3.) * p0 + T(2.) / T(3.) * (ps + dtau * rhs);
}

// RPD planes (and phi0 values) are in flight ahead of the step that uses them.  The
// steps of a chunk run in groups of RPD, unrolled, so the chunk length must be a
// multiple of RPD (the launcher falls back to RPD = 1 otherwise): a step skipped
// inside a group would make the memory-counter waits of every step conservative.
#ifndef ICLS_REINIT_WPE
#define ICLS_REINIT_WPE 1
#endif
template <class T, int STAGE, bool P2, int RPD, bool WALL>
__device__ __forceinline__ void reinit_zm_body(const Grid& g, int bid, int RZ, T NB2, double* res, const T* phi0,
                                               const T* __restrict__ phis, T* out, T* __restrict__ stash) {
    __shared__ T ring[RNS][RTP];
    __shared__ T p0s[RR][BX * BY];
    __shared__ int wok[RNS][BX * BY / 64];   // per plane and wavefront: all values equal wv
    __shared__ T wv[RNS][BX * BY / 64];
    __shared__ double smem[16];
    const int ntx = (g.nx + BX - 1) / BX, nty = (g.ny + RTY - 1) / RTY, ntz = (g.nz + RZ - 1) / RZ;
    const int tile = xcd_tile_dealt(bid, ntx * nty * ntz, XCD_DEAL);
    const int bx = tile % ntx, by = (tile / ntx) % nty, bz = tile / (ntx * nty);
    const int x0 = bx * BX, y0 = by * RTY, kb = 1 + bz * RZ;
    const int ke = kb + RZ - 1 < g.nz ? kb + RZ - 1 : g.nz;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + BX * ty;
    if (STAGE == 1 && tile == 0 && tid == 0) res[0] = 0.;
    const int xm = g.nx - 1;   // nx is a power of two; ny need not be (a window's)
    long poff[RNLD];
#pragma unroll
    for (int r = 0; r < RNLD; r++) {
        int e = tid + r * BX * BY;
        if (e >= RTP) e = RTP - 1;
        const int lx = e % RTW - RH, ly = e / RTW - RH;
        poff[r] = (long)((x0 + lx) & xm) + g.sy * (long)wrap(y0 + ly, g.ny);
    }
    // planes travel global -> registers -> LDS; the register queue holds RPD planes so
    // that a plane's loads are issued RPD steps before the step that stores it
    T nxt[RPD][RNLD];
    auto fetch = [&](int k, T* dst) {
        const long base = g.sz * (long)(k + 2);
#pragma unroll
        for (int r = 0; r < RNLD; r++) dst[r] = phis[base + poff[r]];
    };
    auto slot_of = [&](int k) { return (k - kb + 2 * RNS) & (RNS - 1); };
    auto stash_plane = [&](int k, const T* src) {
        const int sl = slot_of(k);
        T* dst = ring[sl];
        const T v0 = src[0];
        bool ok = true;
#pragma unroll
        for (int r = 0; r < RNLD; r++) {
            const int e = tid + r * BX * BY;
            if (e < RTP) {
                dst[e] = src[r];
                ok = ok && src[r] == v0;
            }
        }
        const T vf = __shfl(v0, 0);
        const int wall = __all(ok && v0 == vf);
        if ((tid & 63) == 0) {
            wok[sl][tid >> 6] = wall;
            wv[sl][tid >> 6] = vf;
        }
    };
    // run: how many of the planes stored so far, the newest last, are one value rval
    // (block-uniform: every wavefront reads the same LDS words after a barrier)
    int run = 0;
    T rval = T(0.);
    auto push = [&](int k) {
        const int sl = slot_of(k);
        const T v = wv[sl][0];
        bool ok = true;
#pragma unroll
        for (int w = 0; w < BX * BY / 64; w++) ok = ok && wok[sl][w] && wv[sl][w] == v;
        if (ok && run > 0 && v == rval) run++;
        else if (ok) { run = 1; rval = v; }
        else run = 0;
    };
    {
        // the six planes before the first step: all loads in flight before the first store
        T pro[2 * RH][RNLD];
#pragma unroll
        for (int t = 0; t < 2 * RH; t++) fetch(kb - RH + t, pro[t]);
#pragma unroll
        for (int t = 0; t < 2 * RH; t++) stash_plane(kb - RH + t, pro[t]);
    }
    __syncthreads();
    for (int k = kb - RH; k < kb + RH; k++) push(k);

    // Thread (tx, ty) computes rows ty, ty + 4, .. (RR of them) of the tile.  A thread
    // past the end of x or y (nx < 64 or ny < 8) takes the wrapped cell and computes
    // and stores the very values of the thread that owns it (its ring positions and
    // loads wrap the same way); no branch guards the arithmetic and the stores, which
    // keeps the memory-counter waits of the loop exact.
    const int i = (x0 + tx) & xm;
    int jr[RR], cr[RR];
#pragma unroll
    for (int rr = 0; rr < RR; rr++) {
        jr[rr] = wrap(y0 + ty + rr * BY, g.ny);
        cr[rr] = (ty + rr * BY + RH) * RTW + tx + RH;
    }
    const T dz = T(g.dz), dtau = T(0.5) * dz;
    const Sp<T, P2> sx{T(g.dx), T(g.dxi)}, sz{T(g.dz), T(g.dzi)};
    T dmax = T(0.);

    // ghost layers (no right-hand side): below the first chunk, above the last; with
    // WALL, layers 0 and k1 get their final values from the wall extrapolation
    T r_lo[RR], r_hi[RR];
#pragma unroll
    for (int rr = 0; rr < RR; rr++) r_lo[rr] = r_hi[rr] = T(0.);
    auto ghost = [&](int k, int j) {
        const long q = gidx(g, i, j, k);
        const T p0 = phi0[q];
        const T r = reinit_combine<T, STAGE, P2>(p0, phis[q], T(0.), dtau);
        if (STAGE == 3) {
            if (k == 0) stash[i + g.sy * j] = p0;
            else if (k == g.nz + 1) stash[g.sz + i + g.sy * j] = p0;
            else dmax = fmax(dmax, fabs(r - p0));
        }
        out[q] = r;
    };
    if (bz == 0)
        for (int k = -2; k <= (WALL ? -1 : 0); k++)
#pragma unroll
            for (int rr = 0; rr < RR; rr++) ghost(k, jr[rr]);

    // The queues: planes of phi_s and the cell's phi0, RPD z steps ahead.  Every queue
    // load is unconditional, its plane clamped to the last one the chunk reads (a
    // conditional load would be merged into its register by a copy, and the copy
    // waits for every load in flight), and the prologue issues them in the order of
    // the loop, so that the waits on the loop's entry path are no tighter than on its
    // back edge.  Each step moves its phi0 to the thread's own LDS slot before the
    // queue register is loaded again, for the same reason.
    T p0q[RPD][RR];
#pragma unroll
    for (int u = 0; u < RPD; u++) {
        fetch(min(kb + u, ke) + RH, nxt[u]);
#pragma unroll
        for (int rr = 0; rr < RR; rr++) p0q[u][rr] = phi0[gidx(g, i, jr[rr], min(kb + u, ke))];
    }
    for (int k0 = kb; k0 <= ke; k0 += RPD) {
#pragma unroll
        for (int u = 0; u < RPD; u++) {
            const int k = k0 + u;
            stash_plane(k + RH, nxt[u]);
#pragma unroll
            for (int rr = 0; rr < RR; rr++) p0s[rr][tid] = p0q[u][rr];
            __syncthreads();
            push(k + RH);
            fetch(min(k + RPD, ke) + RH, nxt[u]);
#pragma unroll
            for (int rr = 0; rr < RR; rr++) p0q[u][rr] = phi0[gidx(g, i, jr[rr], min(k + RPD, ke))];
#pragma unroll
            for (int rr = 0; rr < RR; rr++) {
                const int j = jr[rr], c0 = cr[rr];
                const T p0 = p0s[rr][tid];
                const long q = gidx(g, i, j, k);
                const T* P0 = ring[slot_of(k)];
                const bool flat = run >= 2 * RH + 1;   // planes k - 3 .. k + 3 are all rval
                const T ps = flat ? rval : P0[c0];
                T rhs = T(0.);
                if (flat) {
                    if (fabs(ps) < NB2 + dz) {
                        // six derivatives +0: s * (1 - sqrt(0)) = s
                        if (p0 < -dz) rhs = T(-1.);
                        else if (p0 > dz) rhs = T(1.);
                        else rhs = qdiv(p0, sqrt(p0 * p0 + dz * dz));
                    }
                } else if (fabs(ps) < NB2 + dz) {
                    T s;
                    if (p0 < -dz) s = T(-1.);
                    else if (p0 > dz) s = T(1.);
                    else s = qdiv(p0, sqrt(p0 * p0 + dz * dz));
                    // reinit_rhs on the tile (hyperPDE.f90:2662-2736, 2932-3197)
                    T f[7];
#pragma unroll
                    for (int d = -3; d <= 3; d++) f[d + 3] = P0[c0 + d];
                    T A = T(0.), B = T(0.);
                    if (!flat7(f)) { A = weno_minus(f + 3, sx); B = weno_plus(f + 3, sx); }
#pragma unroll
                    for (int d = -3; d <= 3; d++) f[d + 3] = P0[c0 + d * RTW];
                    T C = T(0.), D = T(0.);
                    if (!flat7(f)) { C = weno_minus(f + 3, sx); D = weno_plus(f + 3, sx); }
                    T E = T(0.), F = T(0.);
                    if (k >= 3 && k <= g.nz - 2) {
#pragma unroll
                        for (int d = -3; d <= 3; d++) f[d + 3] = ring[slot_of(k + d)][c0];
                        if (!flat7(f)) { E = weno_minus(f + 3, sz); F = weno_plus(f + 3, sz); }
                    } else {
                        E = sz.div(ps - ring[slot_of(k - 1)][c0]);
                        F = sz.div(ring[slot_of(k + 1)][c0] - ps);
                    }
                    const T z = T(0.);
                    T dydx, dydy, dydz;
                    if ((B * s < z) && ((A + B) * s < z)) dydx = B;
                    else if ((A * s > z) && ((A + B) * s > z)) dydx = A;
                    else dydx = T(0.5) * (A + B);
                    if ((D * s < z) && ((C + D) * s < z)) dydy = D;
                    else if ((C * s > z) && ((C + D) * s > z)) dydy = C;
                    else dydy = T(0.5) * (C + D);
                    if ((F * s < z) && ((E + F) * s < z)) dydz = F;
                    else if ((E * s > z) && ((E + F) * s > z)) dydz = E;
                    else dydz = T(0.5) * (E + F);
                    rhs = s * (T(1.) - sqrt(dydx * dydx + dydy * dydy + dydz * dydz));
                }
                const T r = reinit_combine<T, STAGE, P2>(p0, ps, rhs, dtau);
                if (STAGE == 3) dmax = fmax(dmax, fabs(r - p0));
                out[q] = r;
                if (WALL) {
                    // layer 0 from layers 1 and 2, layer k1 from kmax - 1 and kmax (k is
                    // uniform over the block)
                    if (k == 1) r_lo[rr] = r;
                    else if (k == 2) {
                        // phi0 of layer 0 is read before the store (stage 3 may run in place)
                        const long qw = gidx(g, i, j, 0);
                        const T lo = T(2.) * r_lo[rr] - r;
                        if (STAGE == 3) dmax = fmax(dmax, fabs(lo - phi0[qw]));
                        out[qw] = lo;
                    }
                    if (k == g.nz - 1) r_hi[rr] = r;
                    else if (k == g.nz) {
                        const T hi = T(2.) * r - r_hi[rr];
                        const long qw = gidx(g, i, j, g.nz + 1);
                        if (STAGE == 3) dmax = fmax(dmax, fabs(hi - phi0[qw]));
                        out[qw] = hi;
                    }
                }
            }
            // no barrier here: the next step stores plane k + 4 into the slot of plane
            // k - 4, which no thread reads after the barrier above
        }
    }
    if (bz == ntz - 1)
        for (int k = g.nz + (WALL ? 2 : 1); k <= g.nz + 3; k++)
#pragma unroll
            for (int rr = 0; rr < RR; rr++) ghost(k, jr[rr]);
    if (STAGE == 3) {
        double b = block_reduce((double)dmax, OpMax(), smem);
        if (tid == 0) atomic_max_nonneg(res, b);
    }
}

template <class T, int STAGE, bool P2, int RPD, bool WALL>
__global__ void __launch_bounds__(BX * BY, ICLS_REINIT_WPE)
k_reinit_zm(Grid g, int RZ, T NB2, double* res, const T* phi0, const T* __restrict__ phis, T* out,
            T* __restrict__ stash) {
    reinit_zm_body<T, STAGE, P2, RPD, WALL>(g, blockIdx.x, RZ, NB2, res, phi0, phis, out, stash);
}

// the same over the windows of a batch (g: the whole grid); window t keeps its
// residual in res[lidx[t]] and its two stash planes at stash + 2 t sz
template <class T, int STAGE, bool P2, int RPD, bool WALL>
__global__ void __launch_bounds__(BX * BY, ICLS_REINIT_WPE)
k_reinit_zm_batch(Grid g, WinBatch wb, int RZ, T NB2, double* res, const T* phi0, const T* phis, T* out,
                  T* stash) {
    const WinPick w = win_pick(g, wb, blockIdx.x);
    reinit_zm_body<T, STAGE, P2, RPD, WALL>(w.g, w.bid, RZ, NB2, res + wb.lidx[w.t], win_ptr(phi0, wb, w),
                                            win_ptr(phis, wb, w), win_ptr(out, wb, w),
                                            stash + 2 * g.sz * (long)w.t + w.poff);
}

template <class T, int RPD, bool WALL>
static void reinit_dispatch_pd(Ctx& c, int rz, int stage, const double* phi0, const double* phis, double* out) {
    dim3 blk(BX, BY, 1);
    const int ntiles = ((c.g.nx + BX - 1) / BX) * ((c.g.ny + RTY - 1) / RTY) * ((c.g.nz + rz - 1) / rz);
    double* res = c.dscal + c.res_slot;
#define RS(S, P)                                                                                                  \
    hipLaunchKernelGGL((k_reinit_zm<T, S, P, RPD, WALL>), dim3(ntiles), blk, 0, c.stream, c.g, rz, (T)c.NB2, res, \
                       as<T>(phi0), as<T>(phis), as<T>(out), as<T>(c.stash))
    if (c.g.pow2) {
        if (stage == 1) RS(1, true); else if (stage == 2) RS(2, true); else RS(3, true);
    } else {
        if (stage == 1) RS(1, false); else if (stage == 2) RS(2, false); else RS(3, false);
    }
#undef RS
}

// Layers per block of the re-init stages: 32.  At 256^3 (MI355X) chunks of 8, 12, 16,
// 24 and 32 layers ran stages 1 / 2-3 in 0.122 / 0.138, 0.130 / 0.140, 0.114 / 0.121,
// 0.139 / 0.151 and 0.098 / 0.105 ms (FP32), and chunks filling whole rounds of
// resident blocks (52 layers) in 0.131 / 0.137 ms: far-field blocks end early (the
// uniform-window path), so block count matters more than rounds.  With the tiles dealt
// to the XCDs in groups (XCD_DEAL), chunks of 8 and 16 layers were again slower than 32.
// Inside a level-set window (a few tens of thousands of cells) 8-layer chunks: four
// times the blocks of 32-layer chunks, and a quarter of each block's serial march.
// With one barrier per step and the prologue loads in flight together, 40, 52 and 64
// layers ran the two FP32 stage 2-3 launches of a step in 0.172, 0.203 and 0.197 ms
// together against 0.169 ms for 32, and the FP64 ones in 0.298, 0.329 and 0.345 ms
// against 0.321 ms: FP64 (three resident blocks per CU instead of five) takes 40.
constexpr int REINIT_RZ = 32, REINIT_RZ_FP64 = 40, REINIT_RZ_WINDOW = 8;
template <class T>
static int reinit_rz(Ctx& c) {
    const int rz = in_window(c) ? REINIT_RZ_WINDOW : (sizeof(T) == 4 ? REINIT_RZ : REINIT_RZ_FP64);
    return c.g.nz < rz ? c.g.nz : rz;
}

// the wall extrapolation folds into the stage when the last chunk has two layers or
// more (a chunk of one layer would need layer kmax - 1 from another block)
static bool fuses_wall(const Grid& g, int rz) { return g.nz >= 2 && (g.nz % rz == 0 || g.nz % rz >= 2); }
static bool reinit_fuses_wall(Ctx& c) {
    return fuses_wall(c.g, c.fp32 ? reinit_rz<float>(c) : reinit_rz<double>(c));
}

template <class T>
static void reinit_dispatch(Ctx& c, int stage, const double* phi0, const double* phis, double* out) {
    const int rz = reinit_rz<T>(c);
    // every chunk (rz layers, the last one nz % rz) a multiple of RPD_MAX
    const bool deep = rz % RPD_MAX == 0 && (c.g.nz % rz) % RPD_MAX == 0;
    if (fuses_wall(c.g, rz)) {
        if (deep) reinit_dispatch_pd<T, RPD_MAX, true>(c, rz, stage, phi0, phis, out);
        else reinit_dispatch_pd<T, 1, true>(c, rz, stage, phi0, phis, out);
    } else {
        if (deep) reinit_dispatch_pd<T, RPD_MAX, false>(c, rz, stage, phi0, phis, out);
        else reinit_dispatch_pd<T, 1, false>(c, rz, stage, phi0, phis, out);
    }
}

// One re-initialisation stage with its boundsls (bound.f90:404-411); after stage 3
// also the wall layers' share of the residual (levelset.f90:831).
int launch_reinit_stage(Ctx& c, int stage, const double* phi0, const double* phis, double* out) {
    {
        ProfScope ps(c, stage == 1 ? "k_reinit_stage1" : "k_reinit_stage23");
        if (c.fp32) reinit_dispatch<float>(c, stage, phi0, phis, out);
        else reinit_dispatch<double>(c, stage, phi0, phis, out);
        HIP_OK(hipGetLastError());
    }
    if (reinit_fuses_wall(c)) return ICLS_OK;
    if (stage == 3) return bc_ls(c, out, c.stash, c.dscal + c.res_slot, nullptr);
    return bc_ls(c, out, nullptr, nullptr, nullptr);
}

// ---------------------------------------------------------------- batched windows
// The windows of one colour in launches of at most WB_MAX windows, each launch holding
// windows with the same key (template and chunk choices).  Chunks of 8 layers, as in a
// one-window launch: on 64 drops at 256^3 (MI355X) chunks of 4, 8, 16 and 32 layers
// gave 8.78, 8.69, 9.03 and 9.97 ms per learned step; the kernels are bound by the
// serial march of each block, so shorter chains beat fewer halo planes.
constexpr int REINIT_RZ_BATCH = 8, TZ_BATCH = 8;
static bool win_full(const Ctx& c, const LsWindow& w) {
    return w.wx == c.gfull.nx && w.wy == c.gfull.ny && w.wz == c.gfull.nz;
}
static Grid win_grid(const Ctx& c, const LsWindow& w) {
    Grid g = c.g;
    g.nx = w.wx; g.ny = w.wy; g.nz = w.wz;
    return g;
}

// tube and the three advection stages of ls_main for the level sets ls, whose windows
// do not overlap: the launches of advect() in icls_runtime, one per colour instead of
// one per level set
template <class T>
static int advect_batch_t(Ctx& c, const std::vector<int>& ls) {
    dim3 blk(BX, BY, 1);
    const T NB2 = (T)c.NB2;
    // tube: chunk length and planes ahead as tube_dispatch chooses them
    auto tz_of = [&](const LsWindow& w) { return win_full(c, w) ? TZ_GRID : TZ_BATCH; };
    auto tkey = [&](const LsWindow& w) {
        const int tz = tz_of(w);
        return tz * 2 + ((tz % 2 == 0 && (w.wz % tz) % 2 == 0) ? 1 : 0);
    };
    auto tblocks = [&](const LsWindow& w) -> long {
        const int tz = tz_of(w);
        return (long)((w.wx + BX - 1) / BX) * ((w.wy + BY - 1) / BY) * ((w.wz + tz - 1) / tz);
    };
    int rc;
    {
        ProfScope ps(c, "k_tube");
        rc = for_each_batch(c, ls, tkey, tblocks, [&](const WinBatch& wb, int k, const std::vector<int>&) -> int {
            const int tz = k / 2;
            if (k & 1)
                hipLaunchKernelGGL((k_tube_zm_batch<T, 2>), dim3(wb.first[wb.n]), blk, 0, c.stream, c.g, wb, tz, NB2,
                                   (const T*)nullptr, as<T>(c.ls_a), c.mask);
            else
                hipLaunchKernelGGL((k_tube_zm_batch<T, 1>), dim3(wb.first[wb.n]), blk, 0, c.stream, c.g, wb, tz, NB2,
                                   (const T*)nullptr, as<T>(c.ls_a), c.mask);
            HIP_OK(hipGetLastError());
            return ICLS_OK;
        });
        if (rc) return rc;
    }
    auto akey = [&](const LsWindow&) { return 0; };
    auto ablocks = [&](const LsWindow& w) -> long {
        return (long)((w.wx + BX - 1) / BX) * ((w.wy + BY - 1) / BY) * (w.wz + 6);
    };
    const T* sa = as<T>(c.ls_a);
    const T* sb = as<T>(c.ls_b);
    const T* sc = as<T>(c.ls_c);
    const T* in[3][2] = {{sa, sa}, {sa, sb}, {sa, sc}};      // phi0, phis per stage
    T* out[3] = {as<T>(c.ls_b), as<T>(c.ls_c), nullptr};     // stage 3 writes the level set
    for (int st = 1; st <= 3; st++) {
        ProfScope ps(c, st == 1 ? "k_adv_stage1" : "k_adv_stage23");
        rc = for_each_batch(c, ls, akey, ablocks, [&](const WinBatch& wb, int, const std::vector<int>&) -> int {
#define ADVB(S, ST, P)                                                                                            \
    hipLaunchKernelGGL((k_adv_stage_batch<T, S, ST, P, true>), dim3(wb.first[wb.n]), blk, 0, c.stream, c.g, wb,     \
                       (T)c.dt, in[ST - 1][0], in[ST - 1][1], out[ST - 1], (const uint8_t*)c.mask, as<T>(c.unew),  \
                       as<T>(c.vnew), as<T>(c.wnew))
#define ADVB_ST(S)                                                                                     \
    do {                                                                                               \
        if (c.g.pow2) { if (st == 1) ADVB(S, 1, true); else if (st == 2) ADVB(S, 2, true); else ADVB(S, 3, true); } \
        else { if (st == 1) ADVB(S, 1, false); else if (st == 2) ADVB(S, 2, false); else ADVB(S, 3, false); }       \
    } while (0)
            if (c.P.ls_adv_space == ICLS_ADV_HOUC5) ADVB_ST(ICLS_ADV_HOUC5);
            else ADVB_ST(ICLS_ADV_WENO5);
#undef ADVB_ST
#undef ADVB
            HIP_OK(hipGetLastError());
            return ICLS_OK;
        });
        if (rc) return rc;
    }
    return ICLS_OK;
}

int launch_advect_batch(Ctx& c, const std::vector<int>& ls) {
    return c.fp32 ? advect_batch_t<float>(c, ls) : advect_batch_t<double>(c, ls);
}

// the three re-initialisation stages for the level sets ls (windows that do not
// overlap), with the chunk length, planes ahead and wall folding of reinit_dispatch
// per window; a window whose last chunk is one layer runs k_bc_ls after each stage
template <class T>
static int reinit_batch_t(Ctx& c, const std::vector<int>& ls) {
    dim3 blk(BX, BY, 1);
    const T NB2 = (T)c.NB2;
    auto rz_of = [&](const LsWindow& w) {
        const int rz = win_full(c, w) ? (sizeof(T) == 4 ? REINIT_RZ : REINIT_RZ_FP64) : REINIT_RZ_BATCH;
        return w.wz < rz ? w.wz : rz;
    };
    // key: rz * 4 + 2 deep + wall
    auto key = [&](const LsWindow& w) {
        const int rz = rz_of(w);
        const bool deep = rz % RPD_MAX == 0 && (w.wz % rz) % RPD_MAX == 0;
        const bool wall = fuses_wall(win_grid(c, w), rz);
        return rz * 4 + (deep ? 2 : 0) + (wall ? 1 : 0);
    };
    auto blocks = [&](const LsWindow& w) -> long {
        const int rz = rz_of(w);
        return (long)((w.wx + BX - 1) / BX) * ((w.wy + RTY - 1) / RTY) * ((w.wz + rz - 1) / rz);
    };
    double* res = c.dscal + c.res_b0;
    T* stash = as<T>(c.stash_b);
    const T* sa = as<T>(c.ls_a);
    const T* sb = as<T>(c.ls_b);
    const T* phis[3] = {nullptr, sa, sb};
    T* out[3] = {as<T>(c.ls_a), as<T>(c.ls_b), nullptr};
    for (int st = 1; st <= 3; st++) {
        ProfScope ps(c, st == 1 ? "k_reinit_stage1" : "k_reinit_stage23");
        int rc = for_each_batch(c, ls, key, blocks, [&](const WinBatch& wb, int k, const std::vector<int>& members) -> int {
            const int rz = k / 4;
            const bool deep = k & 2, wall = k & 1;
#define RB(S, P, PD, W)                                                                                            \
    hipLaunchKernelGGL((k_reinit_zm_batch<T, S, P, PD, W>), dim3(wb.first[wb.n]), blk, 0, c.stream, c.g, wb, rz,   \
                       NB2, res, (const T*)nullptr, phis[S - 1], out[S - 1], stash)
#define RB_S(P, PD, W)                                                                \
    do {                                                                              \
        if (st == 1) RB(1, P, PD, W); else if (st == 2) RB(2, P, PD, W); else RB(3, P, PD, W); \
    } while (0)
#define RB_W(P, PD)                            \
    do {                                       \
        if (wall) RB_S(P, PD, true);           \
        else RB_S(P, PD, false);               \
    } while (0)
            if (c.g.pow2) { if (deep) RB_W(true, RPD_MAX); else RB_W(true, 1); }
            else { if (deep) RB_W(false, RPD_MAX); else RB_W(false, 1); }
#undef RB_W
#undef RB_S
#undef RB
            HIP_OK(hipGetLastError());
            if (!wall) {
                // the wall layers of each window as launch_reinit_stage does without folding
                const Grid g0 = c.g;
                for (int t = 0; t < wb.n; t++) {
                    const int l = members[t];
                    const LsWindow& w = c.win[l];
                    const long poff = w.ox + g0.sy * (long)w.oy, off = poff + g0.sz * (long)w.oz;
                    double* o = st == 3 ? c.lvset[l] : (st == 1 ? c.ls_a : c.ls_b);
                    o = (double*)((char*)o + off * (long)c.esz);
                    const double* ref =
                        st == 3 ? (const double*)((char*)c.stash_b + (2 * g0.sz * (long)t + poff) * (long)c.esz) : nullptr;
                    c.g = win_grid(c, w);
                    int rc2 = bc_ls(c, o, ref, st == 3 ? res + l : nullptr, nullptr);
                    c.g = g0;
                    if (rc2) return rc2;
                }
            }
            if (st == 3)
                for (int t = 0; t < wb.n; t++)
                    if (members[t] == c.lmax - 1)
                        HIP_OK(hipMemcpyAsync(c.dscal + DS_REINIT_RES, res + c.lmax - 1, sizeof(double),
                                              hipMemcpyDeviceToDevice, c.stream));
            return ICLS_OK;
        });
        if (rc) return rc;
    }
    return ICLS_OK;
}

int launch_reinit_batch(Ctx& c, const std::vector<int>& ls) {
    return c.fp32 ? reinit_batch_t<float>(c, ls) : reinit_batch_t<double>(c, ls);
}

// ---------------------------------------------------------------- level-set windows
// Bounding box, in the launch grid's own indices, of the cells with phi < lim (the
// inside and the narrow band of a drop): ext = {min i, min j, min k, -max i, -max j,
// -max k}, folded in with atomicMin (the caller fills ext with a large value).
template <class T>
__global__ void k_ls_extent(Grid g, const T* __restrict__ phi, T lim, int* __restrict__ ext) {
    __shared__ int s[6];
    const int tid = threadIdx.x + BX * threadIdx.y;
    if (tid < 6) s[tid] = 0x7f7f7f7f;
    __syncthreads();
    const int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 < g.nx && j0 < g.ny && phi[gidx(g, i0, j0, k)] < lim) {
        atomicMin(&s[0], i0); atomicMin(&s[1], j0); atomicMin(&s[2], k);
        atomicMin(&s[3], -i0); atomicMin(&s[4], -j0); atomicMin(&s[5], -k);
    }
    __syncthreads();
    if (tid < 6 && s[tid] != 0x7f7f7f7f) atomicMin(&ext[tid], s[tid]);
}

int launch_ls_extent(Ctx& c, const double* phi, int* ext) {
    dim3 blk(BX, BY, 1);
    dim3 grd((c.g.nx + BX - 1) / BX, (c.g.ny + BY - 1) / BY, c.g.nz);
    if (c.fp32)
        hipLaunchKernelGGL(k_ls_extent<float>, grd, blk, 0, c.stream, c.g, as<float>(phi), (float)c.NB2, ext);
    else
        hipLaunchKernelGGL(k_ls_extent<double>, grd, blk, 0, c.stream, c.g, (const double*)phi, c.NB2, ext);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- initial level sets
// Sphere SDF with the +lx and +ly periodic images (init.f90:945-980; 'drp' is the
// lmax = 1 case, init.f90:717-753), evaluated in double at every stored cell and
// stored in the field type.
template <class T>
__global__ void k_init_ls(Grid g, double lx, double ly, double xc, double yc, double zc, double R, T* __restrict__ ph) {
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
    ph[gidx(g, i0, j0, k)] = (T)v;
}

int launch_init_levelset(Ctx& c, int l, double* phi) {
    ProfScope ps(c, "k_init_ls");
    dim3 blk(BX, BY, 1);
    dim3 grd = grid3(c.g, c.g.nz + 6, blk);
    const icls_params& P = c.P;
    if (c.fp32)
        hipLaunchKernelGGL(k_init_ls<float>, grd, blk, 0, c.stream, c.g, P.lx, P.ly, P.drop_xc[l], P.drop_yc[l],
                           P.drop_zc[l], P.drop_r[l], as<float>(phi));
    else
        hipLaunchKernelGGL(k_init_ls<double>, grd, blk, 0, c.stream, c.g, P.lx, P.ly, P.drop_xc[l], P.drop_yc[l],
                           P.drop_zc[l], P.drop_r[l], phi);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

}  // namespace icls
