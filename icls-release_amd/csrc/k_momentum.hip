// k_momentum.hip -- ab2 + mom_main (convection, diffusion, buoyancy, AB2 and the
// split-pressure term), the velocity/pressure boundary conditions, correc, the
// pressure rotation and normalisation, cell-centre velocities and the diagnostic
// reductions (chkdiv, chkdt, max_vel_at_center).
//
// T is the element type of the field arrays; the pressures pnew, pold and p are
// double in every mode, and pressure differences are formed in double before they
// meet a field value.
#include <cstdint>
#include <type_traits>

#include "icls_internal.hpp"
#include "icls_reduce.hpp"
#include "icls_stencil.hpp"

namespace icls {

static constexpr int BX = 64, BY = 4;
static constexpr int MBY = 8;   // tile height of the momentum kernel
static constexpr int KCH = 8;

struct MomParams {
    double dt, visc, iFr, g0, g1, g2, rho0, inv_rho0, rho1, forceytot;
    // variable-step AB2 and p_hat extrapolation with r = dt / dt_old:
    // ab_n = 1 + r/2, ab_o = r/2, ex_n = 1 + r, ex_o = r (1.5, 0.5, 2, 1 at r = 1)
    double ab_n, ab_o, ex_n, ex_o;
    int rho_equal;
    int mz;                     // layers per block (zchunk_for)
    int plazy, pprev, pfirst;   // the lazy pressure rotation (Ctx::p_lazy ..)
    uint32_t fbytes;            // bytes of a field array (with 32-bit offsets: below 2^31)
};

// ab2 (ab2.f90:49-129) with mom_main (mom.f90:13-62), CNTR2_conv (mom.f90:67-130),
// delta_diff (mom.f90:134-273), NS_grav (mom.f90:277-301) and momxp/yp/zp
// (mom.f90:779-851).  Face densities are the arithmetic means of update_materials
// (misc.f90:151-172), formed here from rho.
//
// The kernel marches in z.  A block owns a 64 x 8 tile of (i,j) columns over m.mz
// consecutive layers.  u, v, w, miu and rho -- the arrays read at neighbouring cells
// -- pass through rings of LDS planes of the tile with a one-cell x/y halo, so every
// element of them is fetched once per block.  Step k reads planes k-1 .. k+1 of u, v,
// w and miu and planes k, k+1 of rho; right after the step's barrier the block issues
// the loads of plane k+2 and of the cell's point data for step k+1, computes step k
// while they are in flight, and stores plane k+2 into the slot of a plane no thread
// reads any more (k-2 in the four-slot ring, k-1 in rho's three-slot ring) before the
// one barrier of the step.  The rings of the float instantiation (67 KB with the
// pressure planes) let two blocks of 512 threads share a CU, four waves per SIMD; the
// y halo adds a quarter to the planes read.  Tiles are taken in XCD order (xcd_tile).
// x and y wrap with a mask: itot and jtot are powers of two.
constexpr int TW = BX + 2, TH = MBY + 2, TP = TW * TH;   // tile plane with halo
constexpr int NLD = (TP + BX * MBY - 1) / (BX * MBY);      // plane elements per thread
// pressure planes: the tile with a +1 halo in x and y (the split-pressure term reads
// the cell and its x+1, y+1 neighbours from them, and its z+1 neighbour by a point load)
constexpr int PW = BX + 1, PP = PW * (MBY + 1);
constexpr int PLD = (PP + BX * MBY - 1) / (BX * MBY);

// Element q of a field.  With O32 the index and the byte offset are 32-bit, so an access
// is the field's base (uniform, in scalar registers) plus one 32-bit vector offset and
// needs no 64-bit address arithmetic; the host launches O32 only when every field is
// smaller than 2 GiB.
template <bool O32, class P, class I>
__device__ __forceinline__ P& el(P* b, I q) {
    if constexpr (O32) {
        using C = typename std::conditional<std::is_const<P>::value, const char, char>::type;
        return *reinterpret_cast<P*>(reinterpret_cast<C*>(b) + (uint32_t)q * (uint32_t)sizeof(P));
    } else {
        return b[q];
    }
}

// A buffer resource over the first nbytes bytes of a field (nbytes < 2^31).  A buffer
// load at a byte offset of 2^31 or more is out of range: it returns zero and reads no
// memory, which is how a masked-off face jump is read.
constexpr uint32_t OFF_NONE = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const void* p, uint32_t nbytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)nbytes, 0x00020000);
}
template <class T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    if constexpr (std::is_same<T, float>::value) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    } else {
        const auto d = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
        return __hiloint2double((int)d[1], (int)d[0]);
    }
}

template <class T, int RHS_ONLY, bool P2, bool O32>
__global__ void __launch_bounds__(BX * MBY)
__attribute__((amdgpu_waves_per_eu(std::is_same<T, float>::value && O32 ? 4 : 1)))
k_momentum_zm(Grid g, MomParams m, const double* __restrict__ dscal, const T* __restrict__ u, const T* __restrict__ v,
              const T* __restrict__ w, const T* __restrict__ miu, const T* __restrict__ rho, T* __restrict__ duold,
              T* __restrict__ dvold, T* __restrict__ dwold, const double* __restrict__ pnew,
              const double* __restrict__ pold, const T* __restrict__ p_x, const T* __restrict__ p_y,
              const T* __restrict__ p_z, const T* __restrict__ p_xold, const T* __restrict__ p_yold,
              const T* __restrict__ p_zold, T* __restrict__ dudt, T* __restrict__ dvdt, T* __restrict__ dwdt,
              const uint8_t* __restrict__ jm, const uint8_t* __restrict__ jmo, const T* __restrict__ zp) {
    __shared__ T ring[4][4][TP];   // u, v, w, miu x four planes
    __shared__ T rring[3][TP];     // rho x three planes (read at layers k and k+1 only)
    // pnew and pold x two planes: at step k the derive for layer k+1 reads plane k+1
    // (the cell and its x+1, y+1 neighbours) while plane k+2 is stored into the slot of
    // plane k; the z+1 neighbour is a point load of the cell's own column
    __shared__ double pring[2][2][RHS_ONLY ? 1 : PP];
    using Ix = typename std::conditional<O32, uint32_t, long>::type;
    auto cell = [&](int i_, int j_, int k_) -> Ix { return (Ix)i_ + (Ix)g.sy * (Ix)j_ + (Ix)g.sz * (Ix)(k_ + 2); };
    const int MZ = m.mz;
    const int ntx = (g.nx + BX - 1) / BX, nty = (g.ny + MBY - 1) / MBY, ntz = (g.nz + MZ - 1) / MZ;
    const int tile = xcd_tile(blockIdx.x, ntx * nty * ntz);
    const int bx = tile % ntx, by = (tile / ntx) % nty, bz = tile / (ntx * nty);
    const int x0 = bx * BX, y0 = by * MBY, kb = 1 + bz * MZ;
    const int ke = kb + MZ - 1 < g.nz ? kb + MZ - 1 : g.nz;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + BX * ty;
    const int xm = g.nx - 1, ym = g.ny - 1;
    const T* src[5] = {u, v, w, miu, rho};
    // plane element e of the tile -> global offset within a z layer
    Ix poff[NLD];
#pragma unroll
    for (int r = 0; r < NLD; r++) {
        int e = tid + r * BX * MBY;
        if (e >= TP) e = TP - 1;
        const int lx = e % TW - 1, ly = e / TW - 1;
        poff[r] = (Ix)((x0 + lx) & xm) + (Ix)g.sy * (Ix)((y0 + ly) & ym);
    }
    Ix qoff[PLD];
#pragma unroll
    for (int r = 0; r < PLD; r++) {
        int e = tid + r * BX * MBY;
        if (e >= PP) e = PP - 1;
        qoff[r] = (Ix)((x0 + e % PW) & xm) + (Ix)g.sy * (Ix)((y0 + e / PW) & ym);
    }
    double pnx[2][PLD];
    auto pfetch = [&](int k) {
        if (RHS_ONLY) return;
        const Ix base = (Ix)g.sz * (Ix)(k + 2);
#pragma unroll
        for (int r = 0; r < PLD; r++) {
            pnx[0][r] = el<O32>(pnew, base + qoff[r]);
            pnx[1][r] = el<O32>(pold, base + qoff[r]);
        }
    };
    auto pslot = [&](int k) { return (k - kb) & 1; };   // k >= kb
    auto pstash = [&](int k) {
        if (RHS_ONLY) return;
        const int sl = pslot(k);
#pragma unroll
        for (int r = 0; r < PLD; r++) {
            const int e = tid + r * BX * MBY;
            if (e < PP) {
                pring[0][sl][e] = pnx[0][r];
                pring[1][sl][e] = pnx[1][r];
            }
        }
    };
    // planes travel global -> registers -> LDS
    T nxt[5][NLD];
    auto fetch = [&](int k, T (*dst)[NLD]) {
        const Ix base = (Ix)g.sz * (Ix)(k + 2);
#pragma unroll
        for (int f = 0; f < 5; f++)
#pragma unroll
            for (int r = 0; r < NLD; r++) dst[f][r] = el<O32>(src[f], base + poff[r]);
    };
    auto slot_of = [&](int k) { return (k - kb + 4) & 3; };   // k >= kb - 1
    auto rslot_of = [&](int k) { return (k - kb + 4) % 3; };  // k >= kb - 1
    auto stash = [&](int k, T (*from)[NLD]) {
        const int sl = slot_of(k), rs = rslot_of(k);
#pragma unroll
        for (int r = 0; r < NLD; r++) {
            const int e = tid + r * BX * MBY;
            if (e < TP) {
#pragma unroll
                for (int f = 0; f < 4; f++) ring[f][sl][e] = from[f][r];
                rring[rs][e] = from[4][r];
            }
        }
    };
    if constexpr (std::is_same<T, float>::value) {
        // the three planes before the first step: all loads in flight before the first
        // store (in FP64 the extra registers cost more than the overlap gains)
        T pro[3][5][NLD];
#pragma unroll
        for (int t = 0; t < 3; t++) fetch(kb - 1 + t, pro[t]);
#pragma unroll
        for (int t = 0; t < 3; t++) stash(kb - 1 + t, pro[t]);
    } else {
        for (int k = kb - 1; k <= kb + 1; k++) {
            fetch(k, nxt);
            stash(k, nxt);
        }
    }
    // pressure planes kb and kb + 1 (no further than ke, the last one read)
    for (int k = kb; k <= kb + 1 && k <= ke; k++) {
        pfetch(k);
        pstash(k);
    }

    // A thread past the end of x or y (nx < 64 or ny < 8) takes the wrapped cell: its
    // ring positions and point loads wrap the same way, so it computes and stores the
    // very values of the thread that owns the cell.  No branch guards the arithmetic
    // and the stores, which keeps the memory-counter waits of the loop exact.
    const int i = (x0 + tx) & xm, j = (y0 + ty) & ym;
    // the cell's point data (AB2 history, pressures, face jumps), loaded one z step
    // ahead and turned into the values the step uses (du0 .. gpz) before it is loaded
    // again.  Every load of the loop is unconditional, its layer clamped to the last
    // one the chunk reads: a conditional load would be merged into its register by a
    // copy, and the copy waits for every load in flight.
    // The face jumps are read only where the GFM masks say they may be nonzero; the
    // mask bytes travel one step further ahead (jn, jo: the layer fetch_point loads
    // next), so that the jump loads do not wait for them.  With 32-bit offsets a
    // masked-off lane gives its buffer load an out-of-range offset, which returns zero
    // without reading memory; with 64-bit offsets it loads a zero of the zero plane zp
    // (cached, the same words every layer).  Either way the load stays unconditional,
    // with no branch and no select after it.
    T pd[9];
    double pp[6], pz[2];
    unsigned jn = 7u, jo = 7u;
    const uint32_t fb = m.fbytes;
    const __amdgpu_buffer_rsrc_t rjx = field_rsrc(p_x, fb), rjxo = field_rsrc(p_xold, fb);
    const __amdgpu_buffer_rsrc_t rjy = field_rsrc(p_y, fb), rjyo = field_rsrc(p_yold, fb);
    const __amdgpu_buffer_rsrc_t rjz = field_rsrc(p_z, fb), rjzo = field_rsrc(p_zold, fb);
    auto fetch_mask = [&](int k) {
        if (RHS_ONLY) return;
        const Ix q = cell(i, j, k);
        jn = el<O32>(jm, q);
        jo = el<O32>(jmo, q);
    };
    auto fetch_point = [&](int k) {
        if (RHS_ONLY) return;
        const Ix q = cell(i, j, k);
        pd[0] = el<O32>(duold, q); pd[1] = el<O32>(dvold, q); pd[2] = el<O32>(dwold, q);
        // the pressures of the z+1 neighbour, for the split-pressure term of layer k
        pz[0] = el<O32>(pnew, q + (Ix)g.sz);
        pz[1] = el<O32>(pold, q + (Ix)g.sz);
        if constexpr (O32) {
            const uint32_t b = (uint32_t)q * (uint32_t)sizeof(T);
            pd[3] = bload<T>(rjx, (jn & 1u) ? b : OFF_NONE);
            pd[4] = bload<T>(rjxo, (jo & 1u) ? b : OFF_NONE);
            pd[5] = bload<T>(rjy, (jn & 2u) ? b : OFF_NONE);
            pd[6] = bload<T>(rjyo, (jo & 2u) ? b : OFF_NONE);
            pd[7] = bload<T>(rjz, (jn & 4u) ? b : OFF_NONE);
            pd[8] = bload<T>(rjzo, (jo & 4u) ? b : OFF_NONE);
        } else {
            const T* z0 = zp + (i + g.sy * (long)j);
            pd[3] = *((jn & 1u) ? &el<O32>(p_x, q) : z0);
            pd[4] = *((jo & 1u) ? &el<O32>(p_xold, q) : z0);
            pd[5] = *((jn & 2u) ? &el<O32>(p_y, q) : z0);
            pd[6] = *((jo & 2u) ? &el<O32>(p_yold, q) : z0);
            pd[7] = *((jn & 4u) ? &el<O32>(p_z, q) : z0);
            pd[8] = *((jo & 4u) ? &el<O32>(p_zold, q) : z0);
        }
    };
    T du0 = T(0.), dv0 = T(0.), dw0 = T(0.);
    T gpx = T(0.), gpy = T(0.), gpz = T(0.);   // formed in double, kept in T
    const double exn = m.ex_n, exo = m.ex_o;
    const T exn_ = T(m.ex_n), exo_ = T(m.ex_o);
    // pnew and pold from the raw buffers of a lazy rotation, with the operations of
    // the rotation itself (main.f90:164-183)
    const double SA = dscal[DS_PN_A], SB = dscal[DS_PN_B];
    auto pn_of = [&](double a) -> double { return m.plazy ? a - SA : a; };
    auto po_of = [&](double b) -> double {
        if (!m.plazy) return b;
        const double po = m.pprev ? b - SB : b;
        return m.pfirst ? po - SA : po + SB - SA;
    };
    const int cp = ty * PW + tx;   // own cell in the pressure plane
    auto derive = [&](int L) {
        if (RHS_ONLY) return;
        {
            const double* N0 = pring[0][pslot(L)];
            const double* O0 = pring[1][pslot(L)];
            pp[0] = N0[cp]; pp[1] = O0[cp];
            pp[2] = N0[cp + 1]; pp[3] = O0[cp + 1];
            pp[4] = N0[cp + PW]; pp[5] = O0[cp + PW];
        }
        du0 = pd[0]; dv0 = pd[1]; dw0 = pd[2];
        // split pressure (ab2.f90:79-101): p_hat = 2 p^n - p^(n-1), p_hat_j = 2 p_j - p_j,old
        // (ex_n = 2, ex_o = 1 at a constant dt)
        const T pjx = exn_ * pd[3] - exo_ * pd[4];
        const T pjy = exn_ * pd[5] - exo_ * pd[6];
        const T pjz = exn_ * pd[7] - exo_ * pd[8];
        if constexpr (std::is_same<T, double>::value) {
            // the reference's operations: p_hat at both cells, then the difference
            const double phq = exn * pn_of(pp[0]) - exo * po_of(pp[1]);
            const double phx = exn * pn_of(pp[2]) - exo * po_of(pp[3]);
            const double phy = exn * pn_of(pp[4]) - exo * po_of(pp[5]);
            const double phz = exn * pn_of(pz[0]) - exo * po_of(pz[1]);
            gpx = phx - (double)pjx - phq;
            gpy = phy - (double)pjy - phq;
            gpz = phz - (double)pjz - phq;
        } else {
            // float fields use only the differences p_hat(neighbour) - p_hat(cell), still
            // formed in double; the lazy rotation's offsets are the same at every cell and
            // cancel in them, so the raw buffers are differenced directly
            gpx = T((exn * (pp[2] - pp[0]) - exo * (pp[3] - pp[1])) - (double)pjx);
            gpy = T((exn * (pp[4] - pp[0]) - exo * (pp[5] - pp[1])) - (double)pjy);
            gpz = T((exn * (pz[0] - pp[0]) - exo * (pz[1] - pp[1])) - (double)pjz);
        }
    };
    fetch_mask(kb);
    fetch_point(kb);
    __syncthreads();
    derive(kb);
    fetch_mask(min(kb + 1, ke));
    // the first step stores pressure plane kb + 2 into the slot of plane kb, read just above
    if (!RHS_ONLY) __syncthreads();
    const Sp<T, P2> sx{T(g.dx), T(g.dxi)}, sy{T(g.dy), T(g.dyi)}, sz{T(g.dz), T(g.dzi)};
    const T qt = T(0.25), two = T(2.), one = T(1.);
    const T rho_avr = T(dscal[DS_RHO_AVR]);
    const T iFr = T(m.iFr), visc = T(m.visc), dt = T(m.dt), c15 = T(m.ab_n), c05 = T(m.ab_o), irho0 = T(m.inv_rho0);
    const int c0 = (ty + 1) * TW + tx + 1;   // own cell in the tile plane

    for (int k = kb; k <= ke; k++) {
        const Ix q = cell(i, j, k);
        fetch(min(k + 2, ke + 1), nxt);
        pfetch(min(k + 2, ke));
        fetch_point(min(k + 1, ke));
        fetch_mask(min(k + 2, ke));
        const T* U[3] = {ring[0][slot_of(k - 1)], ring[0][slot_of(k)], ring[0][slot_of(k + 1)]};
        const T* V[3] = {ring[1][slot_of(k - 1)], ring[1][slot_of(k)], ring[1][slot_of(k + 1)]};
        const T* W[3] = {ring[2][slot_of(k - 1)], ring[2][slot_of(k)], ring[2][slot_of(k + 1)]};
        const T* MU[3] = {ring[3][slot_of(k - 1)], ring[3][slot_of(k)], ring[3][slot_of(k + 1)]};
        const T* RH = rring[rslot_of(k)];
        const T* RHP = rring[rslot_of(k + 1)];
        // a(dx, dy, dz) at offset (dx, dy) in plane k + dz
#define AT(A, dx_, dy_, dz_) A[(dz_) + 1][c0 + (dy_) * TW + (dx_)]
        {
            const T uq = AT(U, 0, 0, 0), vq = AT(V, 0, 0, 0), wq = AT(W, 0, 0, 0);
            // ---- convection (CNTR2)
            T uuip = qt * (AT(U, 1, 0, 0) + uq) * (AT(U, 1, 0, 0) + uq);
            T uuim = qt * (AT(U, -1, 0, 0) + uq) * (AT(U, -1, 0, 0) + uq);
            T uvjp = qt * (AT(U, 0, 1, 0) + uq) * (AT(V, 1, 0, 0) + vq);
            T uvjm = qt * (AT(U, 0, -1, 0) + uq) * (AT(V, 1, -1, 0) + AT(V, 0, -1, 0));
            T uwkp = qt * (AT(U, 0, 0, 1) + uq) * (AT(W, 1, 0, 0) + wq);
            T uwkm = qt * (AT(U, 0, 0, -1) + uq) * (AT(W, 1, 0, -1) + AT(W, 0, 0, -1));
            T RCu = sx.div(-uuip + uuim) + sy.div(-uvjp + uvjm) + sz.div(-uwkp + uwkm);
            T uvip = qt * (uq + AT(U, 0, 1, 0)) * (vq + AT(V, 1, 0, 0));
            T uvim = qt * (AT(U, -1, 0, 0) + AT(U, -1, 1, 0)) * (vq + AT(V, -1, 0, 0));
            T vvjp = qt * (vq + AT(V, 0, 1, 0)) * (vq + AT(V, 0, 1, 0));
            T vvjm = qt * (vq + AT(V, 0, -1, 0)) * (vq + AT(V, 0, -1, 0));
            T wvkp = qt * (wq + AT(W, 0, 1, 0)) * (AT(V, 0, 0, 1) + vq);
            T wvkm = qt * (AT(W, 0, 0, -1) + AT(W, 0, 1, -1)) * (AT(V, 0, 0, -1) + vq);
            T RCv = sx.div(-uvip + uvim) + sy.div(-vvjp + vvjm) + sz.div(-wvkp + wvkm);
            T uwip = qt * (wq + AT(W, 1, 0, 0)) * (uq + AT(U, 0, 0, 1));
            T uwim = qt * (wq + AT(W, -1, 0, 0)) * (AT(U, -1, 0, 0) + AT(U, -1, 0, 1));
            T vwjp = qt * (wq + AT(W, 0, 1, 0)) * (vq + AT(V, 0, 0, 1));
            T vwjm = qt * (wq + AT(W, 0, -1, 0)) * (AT(V, 0, -1, 0) + AT(V, 0, -1, 1));
            T wwkp = qt * (wq + AT(W, 0, 0, 1)) * (wq + AT(W, 0, 0, 1));
            T wwkm = qt * (wq + AT(W, 0, 0, -1)) * (wq + AT(W, 0, 0, -1));
            T RCw = sx.div(-uwip + uwim) + sy.div(-vwjp + vwjm) + sz.div(-wwkp + wwkm);

            // ---- face densities
            const T rq = RH[c0];
            T rho_u, rho_v, rho_w;
            if (m.rho_equal) { rho_u = T(m.rho1); rho_v = T(m.rho1); rho_w = T(m.rho1); }
            else {
                rho_u = (rq + RH[c0 + 1]) / two;
                rho_v = (rq + RH[c0 + TW]) / two;
                rho_w = (rq + RHP[c0]) / two;
            }

            // ---- diffusion (delta formulation)
            const T mq = AT(MU, 0, 0, 0);
            T RDu, RDv, RDw;
            {
                T dudxp = sx.div(AT(U, 1, 0, 0) - uq);
                T dudxm = sx.div(uq - AT(U, -1, 0, 0));
                T dudyp = sy.div(AT(U, 0, 1, 0) - uq);
                T dudym = sy.div(uq - AT(U, 0, -1, 0));
                T dvdxp = sx.div(AT(V, 1, 0, 0) - vq);
                T dvdxm = sx.div(AT(V, 1, -1, 0) - AT(V, 0, -1, 0));
                T dudzp = sz.div(AT(U, 0, 0, 1) - uq);
                T dudzm = sz.div(uq - AT(U, 0, 0, -1));
                T dwdxp = sx.div(AT(W, 1, 0, 0) - wq);
                T dwdxm = sx.div(AT(W, 1, 0, -1) - AT(W, 0, 0, -1));
                T ma1 = qt * (mq + AT(MU, 0, -1, 0) + AT(MU, 1, -1, 0) + AT(MU, 1, 0, 0));
                T ma2 = qt * (mq + AT(MU, 0, 1, 0) + AT(MU, 1, 1, 0) + AT(MU, 1, 0, 0));
                T ma3 = qt * (mq + AT(MU, 0, 0, -1) + AT(MU, 1, 0, -1) + AT(MU, 1, 0, 0));
                T ma4 = qt * (mq + AT(MU, 0, 0, 1) + AT(MU, 1, 0, 1) + AT(MU, 1, 0, 0));
                T RD1 = sx.div(two * (dudxp * AT(MU, 1, 0, 0) - dudxm * mq));
                T RD2 = sy.div(dudyp * ma2 - dudym * ma1);
                T RD3 = sy.div(dvdxp * ma2 - dvdxm * ma1);
                T RD4 = sz.div(dudzp * ma4 - dudzm * ma3);
                T RD5 = sz.div(dwdxp * ma4 - dwdxm * ma3);
                RDu = qdiv(RD1 + RD2 + RD3 + RD4 + RD5, rho_u);
            }
            {
                T dvdxp = sx.div(AT(V, 1, 0, 0) - vq);
                T dvdxm = sx.div(vq - AT(V, -1, 0, 0));
                T dudyp = sy.div(AT(U, 0, 1, 0) - uq);
                T dudym = sy.div(AT(U, -1, 1, 0) - AT(U, -1, 0, 0));
                T dvdyp = sy.div(AT(V, 0, 1, 0) - vq);
                T dvdym = sy.div(vq - AT(V, 0, -1, 0));
                T dvdzp = sz.div(AT(V, 0, 0, 1) - vq);
                T dvdzm = sz.div(vq - AT(V, 0, 0, -1));
                T dwdyp = sy.div(AT(W, 0, 1, 0) - wq);
                T dwdym = sy.div(AT(W, 0, 1, -1) - AT(W, 0, 0, -1));
                T ma2 = qt * (mq + AT(MU, 0, 1, 0) + AT(MU, 1, 1, 0) + AT(MU, 1, 0, 0));
                T ma5 = qt * (mq + AT(MU, 0, 1, 0) + AT(MU, -1, 1, 0) + AT(MU, -1, 0, 0));
                T ma6 = qt * (mq + AT(MU, 0, 0, -1) + AT(MU, 0, 1, -1) + AT(MU, 0, 1, 0));
                T ma7 = qt * (mq + AT(MU, 0, 0, 1) + AT(MU, 0, 1, 1) + AT(MU, 0, 1, 0));
                T RD1 = sx.div(dvdxp * ma2 - dvdxm * ma5);
                T RD2 = sx.div(dudyp * ma2 - dudym * ma5);
                T RD3 = sy.div(two * (dvdyp * AT(MU, 0, 1, 0) - dvdym * mq));
                T RD4 = sz.div(dvdzp * ma7 - dvdzm * ma6);
                T RD5 = sz.div(dwdyp * ma7 - dwdym * ma6);
                RDv = qdiv(RD1 + RD2 + RD3 + RD4 + RD5, rho_v);
            }
            {
                T dwdxp = sx.div(AT(W, 1, 0, 0) - wq);
                T dwdxm = sx.div(wq - AT(W, -1, 0, 0));
                T dudzp = sz.div(AT(U, 0, 0, 1) - uq);
                T dudzm = sz.div(AT(U, -1, 0, 1) - AT(U, -1, 0, 0));
                T dwdyp = sy.div(AT(W, 0, 1, 0) - wq);
                T dwdym = sy.div(wq - AT(W, 0, -1, 0));
                T dvdzp = sz.div(AT(V, 0, 0, 1) - vq);
                T dvdzm = sz.div(AT(V, 0, -1, 1) - AT(V, 0, -1, 0));
                T dwdzp = sz.div(AT(W, 0, 0, 1) - wq);
                T dwdzm = sz.div(wq - AT(W, 0, 0, -1));
                T ma4 = qt * (mq + AT(MU, 0, 0, 1) + AT(MU, 1, 0, 1) + AT(MU, 1, 0, 0));
                T ma8 = qt * (mq + AT(MU, 0, 0, 1) + AT(MU, -1, 0, 1) + AT(MU, -1, 0, 0));
                T ma7 = qt * (mq + AT(MU, 0, 0, 1) + AT(MU, 0, 1, 1) + AT(MU, 0, 1, 0));
                T ma9 = qt * (mq + AT(MU, 0, 0, 1) + AT(MU, 0, -1, 1) + AT(MU, 0, -1, 0));
                T RD1 = sx.div(dwdxp * ma4 - dwdxm * ma8);
                T RD2 = sx.div(dudzp * ma4 - dudzm * ma8);
                T RD3 = sy.div(dwdyp * ma7 - dwdym * ma9);
                T RD4 = sy.div(dvdzp * ma7 - dvdzm * ma9);
                T RD5 = sz.div(two * (dwdzp * AT(MU, 0, 0, 1) - dwdzm * mq));
                RDw = qdiv(RD1 + RD2 + RD3 + RD4 + RD5, rho_w);
            }

            // ---- buoyancy and RHS
            T RGu = iFr * (one - qdiv(rho_avr, rho_u)) * T(m.g0);
            T RGv = iFr * (one - qdiv(rho_avr, rho_v)) * T(m.g1);
            T RGw = iFr * (one - qdiv(rho_avr, rho_w)) * T(m.g2);
            T dun = RCu + RDu * visc + RGu;
            T dvn = RCv + RDv * visc + RGv;
            T dwn = RCw + RDw * visc + RGw;
            if (RHS_ONLY) {
                el<O32>(duold, q) = dun; el<O32>(dvold, q) = dvn; el<O32>(dwold, q) = dwn;
            } else {
                // ---- AB2 (ab2.f90:55-65)
                T us = uq + dt * (c15 * dun - c05 * du0);
                T vs = vq + dt * (c15 * dvn - c05 * dv0);
                T ws = wq + dt * (c15 * dwn - c05 * dw0);
                el<O32>(duold, q) = dun; el<O32>(dvold, q) = dvn; el<O32>(dwold, q) = dwn;
                // ---- split-pressure terms (mom.f90:793,818,843)
                T mpx = -(qdiv(one, rho_u) - irho0) * T(g.dxi) * gpx;
                T mpy = -(qdiv(one, rho_v) - irho0) * T(g.dyi) * gpy;
                T mpz = -(qdiv(one, rho_w) - irho0) * T(g.dzi) * gpz;
                el<O32>(dudt, q) = us + dt * mpx;
                el<O32>(dvdt, q) = vs + dt * mpy - qdiv(T(m.forceytot) * dt, rho_v);
                el<O32>(dwdt, q) = ws + dt * mpz;
            }
        }
#undef AT
        derive(min(k + 1, ke));
        stash(k + 2, nxt);
        if (k + 2 <= ke) pstash(k + 2);
        __syncthreads();
    }
}

template <class T>
static void momentum_dispatch(Ctx& c, MomParams m, int rhs_only) {
    dim3 blk(BX, MBY, 1);
    const int txy = ((c.g.nx + BX - 1) / BX) * ((c.g.ny + MBY - 1) / MBY);
    int ntiles = 0;
#define MOM(R, P, O)                                                                                               \
    do {                                                                                                           \
        m.mz = zchunk_for(c, (const void*)k_momentum_zm<T, R, P, O>, BX * MBY, txy, 1, 2);                        \
        ntiles = txy * ((c.g.nz + m.mz - 1) / m.mz);                                                               \
        hipLaunchKernelGGL((k_momentum_zm<T, R, P, O>), dim3(ntiles), blk, 0, c.stream, c.g, m,                     \
                           (const double*)c.dscal, as<T>(c.unew), as<T>(c.vnew), as<T>(c.wnew), as<T>(c.miu),      \
                           as<T>(c.rho), as<T>(c.duold), as<T>(c.dvold), as<T>(c.dwold), (const double*)c.pnew,    \
                           (const double*)c.pold, as<T>(c.p_x), as<T>(c.p_y), as<T>(c.p_z), as<T>(c.p_xold),      \
                           as<T>(c.p_yold), as<T>(c.p_zold), as<T>(c.dudt), as<T>(c.dvdt), as<T>(c.dwdt),    \
                           (const uint8_t*)c.jmask, (const uint8_t*)c.jmask_old, as<T>(c.zplane));                 \
    } while (0)
    // 32-bit offsets when the largest field (a double pressure) is below 2 GiB, so that
    // every offset, and the out-of-range offset 2^31 of a masked-off jump, fit in 32
    // bits; the right-hand-side-only launch (once, at init) keeps 64-bit offsets
    const bool o32 = (unsigned long)c.nfield * sizeof(double) < (1ul << 31);
    m.fbytes = o32 ? (uint32_t)(c.nfield * sizeof(T)) : 0u;
    if (c.g.pow2) {
        if (rhs_only) MOM(1, true, false);
        else if (o32) MOM(0, true, true);
        else MOM(0, true, false);
    } else {
        if (rhs_only) MOM(1, false, false);
        else if (o32) MOM(0, false, true);
        else MOM(0, false, false);
    }
#undef MOM
}

static int momentum_launch(Ctx& c, int rhs_only) {
    ProfScope ps(c, "k_momentum");
    MomParams m;
    m.dt = c.dt; m.visc = c.visc; m.iFr = c.P.iFr;
    const double r = c.dt_old > 0. ? c.dt / c.dt_old : 1.;
    m.ab_n = 1. + r / 2.; m.ab_o = r / 2.;
    m.ex_n = 1. + r; m.ex_o = r;
    if (!rhs_only) c.dt_old = c.dt;      // the next ab2 extrapolates from this step
    m.g0 = c.P.g_unit[0]; m.g1 = c.P.g_unit[1]; m.g2 = c.P.g_unit[2];
    m.rho0 = c.rho0; m.inv_rho0 = 1. / c.rho0; m.rho1 = c.P.rho1;
    m.forceytot = c.P.constPoi ? -12. * c.visc / (c.P.lz * c.P.lz) : 0.;
    m.rho_equal = c.P.rho1 == c.P.rho2;
    m.plazy = c.p_lazy;
    m.pprev = c.p_prev_lazy;
    m.pfirst = c.p_first;
    if (c.fp32) momentum_dispatch<float>(c, m, rhs_only);
    else momentum_dispatch<double>(c, m, rhs_only);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}
int launch_momentum(Ctx& c) { return momentum_launch(c, 0); }
int launch_momentum_rhs_only(Ctx& c) { return momentum_launch(c, 1); }

// ---------------------------------------------------------------- boundary conditions
// bounduvw, BC_in_z = 'Dirichl' (bound.f90:26-47); x/y by wrap
template <class T>
__global__ void k_bc_uvw(Grid g, T vlo, T vup, T* __restrict__ u, T* __restrict__ v, T* __restrict__ w,
                         const double* __restrict__ flag) {
    if (flag && flag[0] == 0.) return;
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    if (i0 >= g.nx || j0 >= g.ny) return;
    const int kmax = g.nz, k1 = g.nz + 1;
    auto I = [&](int k) { return gidx(g, i0, j0, k); };
    u[I(0)] = -u[I(1)];
    u[I(k1)] = -u[I(kmax)];
    v[I(0)] = T(2.) * vlo - v[I(1)];
    v[I(k1)] = T(2.) * vup - v[I(kmax)];
    w[I(0)] = T(0.);
    w[I(kmax)] = T(0.);
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
    if (c.fp32)
        hipLaunchKernelGGL(k_bc_uvw<float>, grid3(c.g, 1, blk), blk, 0, c.stream, c.g, (float)vlo, (float)vup,
                           as<float>(u), as<float>(v), as<float>(w), flag);
    else
        hipLaunchKernelGGL(k_bc_uvw<double>, grid3(c.g, 1, blk), blk, 0, c.stream, c.g, vlo, vup, u, v, w, flag);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}
int launch_bc_uvw(Ctx& c, double* u, double* v, double* w) { return launch_bc_uvw_pred(c, u, v, w, nullptr); }

// boundp z part: Neumann (bound.f90:193-198)
template <class T>
__global__ void k_bc_p(Grid g, T* __restrict__ p, const double* __restrict__ flag) {
    if (flag && flag[0] == 0.) return;
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    if (i0 >= g.nx || j0 >= g.ny) return;
    p[gidx(g, i0, j0, 0)] = p[gidx(g, i0, j0, 1)];
    p[gidx(g, i0, j0, g.nz + 1)] = p[gidx(g, i0, j0, g.nz)];
}

// on a pressure-class (double) array
int launch_bc_p(Ctx& c, double* p) {
    ProfScope ps(c, "k_bc_p");
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_bc_p<double>, grid3(c.g, 1, blk), blk, 0, c.stream, c.g, p, (const double*)nullptr);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// on a field-class array (the Heaviside of the mass correction), predicated
int launch_bc_p_t_pred(Ctx& c, double* p, const double* flag) {
    ProfScope ps(c, "k_bc_p");
    dim3 blk(BX, BY, 1);
    if (c.fp32) hipLaunchKernelGGL(k_bc_p<float>, grid3(c.g, 1, blk), blk, 0, c.stream, c.g, as<float>(p), flag);
    else hipLaunchKernelGGL(k_bc_p<double>, grid3(c.g, 1, blk), blk, 0, c.stream, c.g, p, flag);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- correc
// correc.f90:33-35; the pressure difference p(+1) - p - p_x is formed in double.
// A thread takes KCH layers of its column.  The block also sums the new v over its
// cells into part[block] (the interior sum of average_stream_vel, misc.f90:631-656,
// which the next step reads before anything changes v).
template <class T>
__global__ void k_correc(Grid g, T dt, T rho0, const double* __restrict__ p, const T* __restrict__ px,
                         const T* __restrict__ py, const T* __restrict__ pz, const T* __restrict__ du,
                         const T* __restrict__ dv, const T* __restrict__ dw, T* __restrict__ u, T* __restrict__ v,
                         T* __restrict__ w, double* __restrict__ part, const uint8_t* __restrict__ jm,
                         const T* __restrict__ zp) {
    __shared__ double smem[16];
    const int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    const bool valid = i0 < g.nx && j0 < g.ny;
    double vs = 0.;
    if (valid) {
        const int ip = wrap(i0 + 1, g.nx), jp = wrap(j0 + 1, g.ny);
#pragma unroll
        for (int kk = 0; kk < KCH; kk++) {
            const int k = 1 + blockIdx.z * KCH + kk;
            if (k > g.nz) break;
            const long q = gidx(g, i0, j0, k);
            const double pq = p[q];
            // face jumps only where the GFM mask says they may be nonzero
            const unsigned b = jm[q];
            const T* z0 = zp + (i0 + g.sy * (long)j0);
            const T jx = *((b & 1u) ? px + q : z0), jy = *((b & 2u) ? py + q : z0), jz = *((b & 4u) ? pz + q : z0);
            u[q] = du[q] - dt * T(g.dxi) * T(p[gidx(g, ip, j0, k)] - pq - (double)jx) / rho0;
            const T vq = dv[q] - dt * T(g.dyi) * T(p[gidx(g, i0, jp, k)] - pq - (double)jy) / rho0;
            v[q] = vq;
            vs += (double)vq;
            w[q] = dw[q] - dt * T(g.dzi) * T(p[gidx(g, i0, j0, k + 1)] - pq - (double)jz) / rho0;
        }
    }
    const double r = block_reduce(vs, OpSum(), smem);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[block_linear()] = r;
}

int launch_correc(Ctx& c) {
    {
        ProfScope ps(c, "k_correc");
        dim3 blk(BX, BY, 1);
        dim3 grd((c.g.nx + BX - 1) / BX, (c.g.ny + BY - 1) / BY, (c.g.nz + KCH - 1) / KCH);
        if (c.fp32)
            hipLaunchKernelGGL(k_correc<float>, grd, blk, 0, c.stream, c.g, (float)c.dt, (float)c.rho0,
                               (const double*)c.p, as<float>(c.p_x), as<float>(c.p_y), as<float>(c.p_z),
                               as<float>(c.dudt), as<float>(c.dvdt), as<float>(c.dwdt), as<float>(c.unew),
                               as<float>(c.vnew), as<float>(c.wnew), c.dpart, (const uint8_t*)c.jmask,
                               as<float>(c.zplane));
        else
            hipLaunchKernelGGL(k_correc<double>, grd, blk, 0, c.stream, c.g, c.dt, c.rho0, (const double*)c.p,
                               (const double*)c.p_x, (const double*)c.p_y, (const double*)c.p_z,
                               (const double*)c.dudt, (const double*)c.dvdt, (const double*)c.dwdt, c.unew, c.vnew,
                               c.wnew, c.dpart, (const uint8_t*)c.jmask, (const double*)c.zplane);
        HIP_OK(hipGetLastError());
        const int nb = (int)(grd.x * grd.y * grd.z);
        hipLaunchKernelGGL(k_reduce_final<OpSum>, dim3(1), dim3(RED_FINAL_THREADS), 0, c.stream, (const double*)c.dpart, nb, 0.,
                           c.dscal + DS_V_BULK);
        HIP_OK(hipGetLastError());
    }
    c.vbulk_fresh = true;
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

// The rotation pold <- pnew, pnew <- p, both shifted by the norm (main.f90:164-183),
// is a relabelling of the three buffers: the pnew buffer takes the raw solution, the
// pold buffer the previous pnew buffer, and the old pold buffer becomes the next
// solve's work array.  The shifts are recorded (S_A = norm, S_B = norm_old) and
// applied where the values are read (Ctx::p_lazy).
__global__ void k_pressure_lazy_scalars(double* dscal) {
    if (threadIdx.x == 0) {
        dscal[DS_PN_B] = dscal[DS_NORM_OLD];
        dscal[DS_PN_A] = dscal[DS_NORM];
        dscal[DS_NORM_OLD] = dscal[DS_NORM];
    }
}

static void rotate_buffers(Ctx& c, int first_step) {
    double* a = c.p;
    c.p = c.pold;
    c.pold = c.pnew;
    c.pnew = a;
    if (!c.fp32) c.frhs = c.p;   // the FP64 solve works in place on p
    c.p_prev_lazy = c.p_lazy;
    c.p_lazy = true;
    c.p_first = first_step != 0;
    c.p_in_pnew = true;
}

int launch_pressure_rotate(Ctx& c, int first_step) {
    hipLaunchKernelGGL(k_pressure_lazy_scalars, dim3(1), dim3(64), 0, c.stream, c.dscal);
    HIP_OK(hipGetLastError());
    rotate_buffers(c, first_step);
    return ICLS_OK;
}

// The end of a time step in one single-block launch after k_plane_sum: the norm as
// k_norm_final folds it (the same order), the lazy rotation's scalars as
// k_pressure_lazy_scalars sets them, and the NaN flag of k_nan_check on the new v.
template <class T>
__global__ void k_step_end(Grid g, const double* __restrict__ part, int n, double denom, double* dscal,
                           const T* __restrict__ v) {
    __shared__ double smem[16];
    double acc = 0.;
    for (int q = threadIdx.x; q < n; q += blockDim.x) acc += part[q];
    const double r = block_reduce(acc, OpSum(), smem);
    if (threadIdx.x == 0) {
        const double norm = r / denom;
        dscal[DS_NORM] = norm;
        dscal[DS_PN_B] = dscal[DS_NORM_OLD];
        dscal[DS_PN_A] = norm;
        dscal[DS_NORM_OLD] = norm;
        if (isnan(v[gidx(g, 0, 0, 1)])) dscal[DS_NAN_FLAG] = 1.;
    }
}

int launch_step_end(Ctx& c, int first_step) {
    ProfScope ps(c, "k_plane_sum");
    const int nb = c.g.ny + 2;
    hipLaunchKernelGGL(k_plane_sum, dim3(nb), dim3(256), 0, c.stream, c.g, (const double*)c.p, c.dpart);
    if (c.fp32)
        hipLaunchKernelGGL(k_step_end<float>, dim3(1), dim3(256), 0, c.stream, c.g, (const double*)c.dpart, nb,
                           1. * c.g.nx * c.g.ny, c.dscal, as<float>(c.vnew));
    else
        hipLaunchKernelGGL(k_step_end<double>, dim3(1), dim3(256), 0, c.stream, c.g, (const double*)c.dpart, nb,
                           1. * c.g.nx * c.g.ny, c.dscal, (const double*)c.vnew);
    HIP_OK(hipGetLastError());
    rotate_buffers(c, first_step);
    return ICLS_OK;
}

// the values a lazy rotation stands for, written into the pnew and pold buffers over
// k = 0 .. k1 (the layers the reference rotates); pold from its own buffer first
__global__ void k_pressure_materialize(Grid g, int pprev, int pfirst, const double* __restrict__ dscal,
                                       double* __restrict__ A, double* __restrict__ B) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i0 >= g.nx || j0 >= g.ny) return;
    const long q = gidx(g, i0, j0, k);
    const double SA = dscal[DS_PN_A], SB = dscal[DS_PN_B];
    const double b = B[q];
    const double po = pprev ? b - SB : b;
    B[q] = pfirst ? po - SA : po + SB - SA;
    A[q] = A[q] - SA;
}

// Turns a lazy rotation into stored values (for downloads, uploads and boundary
// calls on the pressures).  The raw solution moves to the work array first while
// it is still the reference's p.
int launch_pressure_release_p(Ctx& c) {
    if (c.p_lazy && c.p_in_pnew) {
        HIP_OK(hipMemcpyAsync(c.p, c.pnew, sizeof(double) * c.nfield, hipMemcpyDeviceToDevice, c.stream));
        c.p_in_pnew = false;
    }
    return ICLS_OK;
}

int launch_pressure_materialize(Ctx& c) {
    if (!c.p_lazy) return ICLS_OK;
    if (int rc = launch_pressure_release_p(c)) return rc;
    dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_pressure_materialize, grid3(c.g, c.g.nz + 2, blk), blk, 0, c.stream, c.g,
                       (int)c.p_prev_lazy, (int)c.p_first, (const double*)c.dscal, c.pnew, c.pold);
    HIP_OK(hipGetLastError());
    c.p_lazy = false;
    return ICLS_OK;
}

// ---------------------------------------------------------------- cell-centre velocity
// average_vel_at_center (misc.f90:587-604)
template <class T>
__global__ void k_avg_vel(Grid g, const T* __restrict__ u, const T* __restrict__ v, const T* __restrict__ w,
                          T* __restrict__ uc, T* __restrict__ vc, T* __restrict__ wc) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    long q = gidx(g, i0, j0, k);
    uc[q] = (u[gidx(g, wrap(i0 - 1, g.nx), j0, k)] + u[q]) / T(2.);
    vc[q] = (v[gidx(g, i0, wrap(j0 - 1, g.ny), k)] + v[q]) / T(2.);
    wc[q] = (w[gidx(g, i0, j0, k - 1)] + w[q]) / T(2.);
}

int launch_avg_vel(Ctx& c) {
    ProfScope ps(c, "k_avg_vel");
    dim3 blk(BX, BY, 1);
    dim3 grd = grid3(c.g, c.g.nz, blk);
    if (c.fp32)
        hipLaunchKernelGGL(k_avg_vel<float>, grd, blk, 0, c.stream, c.g, as<float>(c.unew), as<float>(c.vnew),
                           as<float>(c.wnew), as<float>(c.ucc), as<float>(c.vcc), as<float>(c.wcc));
    else
        hipLaunchKernelGGL(k_avg_vel<double>, grd, blk, 0, c.stream, c.g, (const double*)c.unew, (const double*)c.vnew,
                           (const double*)c.wnew, c.ucc, c.vcc, c.wcc);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- NaN and diagnostics
// isnan(vnew(1,1,1)) (main.f90:207-210) into a sticky flag
template <class T>
__global__ void k_nan_check(Grid g, const T* __restrict__ v, double* dscal) {
    if (threadIdx.x != 0) return;
    if (isnan(v[gidx(g, 0, 0, 1)])) dscal[DS_NAN_FLAG] = 1.;
}

int launch_nan_check(Ctx& c) {
    if (c.fp32) hipLaunchKernelGGL(k_nan_check<float>, dim3(1), dim3(64), 0, c.stream, c.g, as<float>(c.vnew), c.dscal);
    else hipLaunchKernelGGL(k_nan_check<double>, dim3(1), dim3(64), 0, c.stream, c.g, (const double*)c.vnew, c.dscal);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

static inline dim3 red_grid(const Grid& g, int nk) {
    return dim3((g.nx + BX - 1) / BX, (g.ny + BY - 1) / BY, (nk + KCH - 1) / KCH);
}

// chkdiv (chkdiv.f90:28-48): sum of div and max |div| with its first location in
// (k, j, i) loop order; partials [sum | max | location].  The divergence is formed in
// double from the stored velocities.
template <class T>
__global__ void k_chkdiv(Grid g, const T* __restrict__ u, const T* __restrict__ v, const T* __restrict__ w,
                         double* __restrict__ part, int nb) {
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
            double div = ((double)w[q] - (double)w[gidx(g, i0, j0, k - 1)]) * g.dzi +
                         ((double)v[q] - (double)v[gidx(g, i0, wrap(j0 - 1, g.ny), k)]) * g.dyi +
                         ((double)u[q] - (double)u[gidx(g, wrap(i0 - 1, g.nx), j0, k)]) * g.dxi;
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

__global__ void k_chkdiv_final(const double* __restrict__ part, int nb, double* tot, double* mx, double* loc) {
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
        *tot = r; *mx = bm; *loc = bl;
    }
}

int launch_chkdiv(Ctx& c) {
    ProfScope ps(c, "k_chkdiv");
    dim3 blk(BX, BY, 1);
    dim3 grd = red_grid(c.g, c.g.nz);
    int nb = grd.x * grd.y * grd.z;
    if (c.fp32)
        hipLaunchKernelGGL(k_chkdiv<float>, grd, blk, 0, c.stream, c.g, as<float>(c.unew), as<float>(c.vnew),
                           as<float>(c.wnew), c.dpart, nb);
    else
        hipLaunchKernelGGL(k_chkdiv<double>, grd, blk, 0, c.stream, c.g, (const double*)c.unew, (const double*)c.vnew,
                           (const double*)c.wnew, c.dpart, nb);
    hipLaunchKernelGGL(k_chkdiv_final, dim3(1), dim3(256), 0, c.stream, (const double*)c.dpart, nb,
                       c.dscal + DS_DIV_TOT, c.dscal + DS_DIV_MAX, c.dscal + DS_DIV_MAX + 16);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- chkdiv without allocation
// The block partials [sum | max | location] of icls_chkdiv_noalloc live in one fixed
// device array: up to CHKDIV_NOALLOC_BLOCKS reduction blocks, which a 512^3 grid fills
// (8 x 128 x 64 blocks of 64 x 4 columns over 8 layers); 1.5 MiB per GPU, held until
// the process ends.  Every call on a GPU shares that GPU's copy, so calls that overlap
// on one GPU race.  Each block writes its own three partials before the final kernel
// reads them, so nothing has to start at zero.
constexpr int CHKDIV_NOALLOC_BLOCKS = 1 << 16;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[3 * CHKDIV_NOALLOC_BLOCKS];
// The device address of `scratch` on each GPU (host code must not pass the symbol itself).
static double* g_scratch_dev[NOALLOC_MAX_DEVICES];

// Looks the array up once per GPU, with that GPU current.  Called by icls_init for its
// device and by the first icls_chkdiv_noalloc call on any other device.
int noalloc_setup(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) {
        set_error("device %d beyond the %d the allocation-free entry supports", device, NOALLOC_MAX_DEVICES);
        return ICLS_ERR_PARAM;
    }
    if (g_scratch_dev[device]) return ICLS_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        set_error("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        return ICLS_ERR_HIP;
    }
    g_scratch_dev[device] = (double*)ptr;
    return ICLS_OK;
}

// chkdiv (chkdiv.f90:28-48) over caller-owned device arrays in the library's layout,
// the same kernels as launch_chkdiv with the partials in `scratch`; out[0..2] =
// divtot, max |div|, linear index i0 + nx (j0 + ny (k-1)) of its first cell.
int chkdiv_noalloc(const void* u, const void* v, const void* w, int elem_bytes, const Grid& g, double* out,
                   hipStream_t stream) {
    int dev = -1;
    HIP_OK(hipGetDevice(&dev));
    const int rc = noalloc_setup(dev);
    if (rc) return rc;
    const dim3 blk(BX, BY, 1);
    const dim3 grd = red_grid(g, g.nz);
    const long nb = (long)grd.x * grd.y * grd.z;
    if (nb > CHKDIV_NOALLOC_BLOCKS) {
        set_error("grid %d x %d x %d needs %ld reduction blocks; the allocation-free chkdiv holds %d", g.nx, g.ny,
                  g.nz, nb, CHKDIV_NOALLOC_BLOCKS);
        return ICLS_ERR_PARAM;
    }
    double* part = g_scratch_dev[dev];
    if (elem_bytes == 4)
        hipLaunchKernelGGL(k_chkdiv<float>, grd, blk, 0, stream, g, (const float*)u, (const float*)v,
                           (const float*)w, part, (int)nb);
    else
        hipLaunchKernelGGL(k_chkdiv<double>, grd, blk, 0, stream, g, (const double*)u, (const double*)v,
                           (const double*)w, part, (int)nb);
    hipLaunchKernelGGL(k_chkdiv_final, dim3(1), dim3(256), 0, stream, (const double*)part, (int)nb, out, out + 1,
                       out + 2);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// chkdt (chkdt.f90:38-52): max |face-to-centre average| over 1..i1, 1..j1, 1..k1
template <class T>
__global__ void k_umax(Grid g, const T* __restrict__ u, const T* __restrict__ v, const T* __restrict__ w,
                       double* __restrict__ part, int nb) {
    __shared__ double smem[16];
    int i = blockIdx.x * BX + threadIdx.x + 1, j = blockIdx.y * BY + threadIdx.y + 1;   // Fortran 1 .. i1
    T mu = T(0.), mv = T(0.), mw = T(0.);
    if (i <= g.nx + 1 && j <= g.ny + 1) {
        int i0 = wrap(i - 1, g.nx), j0 = wrap(j - 1, g.ny), im = wrap(i - 2, g.nx), jm = wrap(j - 2, g.ny);
        for (int kk = 0; kk < KCH; kk++) {
            int k = 1 + blockIdx.z * KCH + kk;
            if (k > g.nz + 1) break;
            long q = gidx(g, i0, j0, k);
            mu = fmax(mu, fabs((u[q] + u[gidx(g, im, j0, k)]) / T(2.)));
            mv = fmax(mv, fabs((v[q] + v[gidx(g, i0, jm, k)]) / T(2.)));
            mw = fmax(mw, fabs((w[q] + w[gidx(g, i0, j0, k - 1)]) / T(2.)));
        }
    }
    double a = block_reduce((double)mu, OpMax(), smem);
    double b = block_reduce((double)mv, OpMax(), smem);
    double c2 = block_reduce((double)mw, OpMax(), smem);
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
    if (c.fp32)
        hipLaunchKernelGGL(k_umax<float>, grd, blk, 0, c.stream, c.g, as<float>(c.unew), as<float>(c.vnew),
                           as<float>(c.wnew), c.dpart, nb);
    else
        hipLaunchKernelGGL(k_umax<double>, grd, blk, 0, c.stream, c.g, (const double*)c.unew, (const double*)c.vnew,
                           (const double*)c.wnew, c.dpart, nb);
    for (int t = 0; t < 3; t++)
        hipLaunchKernelGGL(k_reduce_final<OpMax>, dim3(1), dim3(RED_FINAL_THREADS), 0, c.stream, (const double*)(c.dpart + t * nb), nb,
                           0., c.dscal + DS_UMAX + t);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// max_vel_at_center (misc.f90:621-632)
template <class T>
__global__ void k_vmag_max(Grid g, const T* __restrict__ uc, const T* __restrict__ vc, const T* __restrict__ wc,
                           double* __restrict__ part) {
    __shared__ double smem[16];
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y;
    T m = T(0.);
    if (i0 < g.nx && j0 < g.ny)
        for (int kk = 0; kk < KCH; kk++) {
            int k = 1 + blockIdx.z * KCH + kk;
            if (k > g.nz) break;
            long q = gidx(g, i0, j0, k);
            m = fmax(m, sqrt(uc[q] * uc[q] + vc[q] * vc[q] + wc[q] * wc[q]));
        }
    double r = block_reduce((double)m, OpMax(), smem);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[block_linear()] = r;
}

int launch_max_vel(Ctx& c) {
    ProfScope ps(c, "k_vmag_max");
    dim3 blk(BX, BY, 1);
    dim3 grd = red_grid(c.g, c.g.nz);
    int nb = grd.x * grd.y * grd.z;
    if (c.fp32)
        hipLaunchKernelGGL(k_vmag_max<float>, grd, blk, 0, c.stream, c.g, as<float>(c.ucc), as<float>(c.vcc),
                           as<float>(c.wcc), c.dpart);
    else
        hipLaunchKernelGGL(k_vmag_max<double>, grd, blk, 0, c.stream, c.g, (const double*)c.ucc, (const double*)c.vcc,
                           (const double*)c.wcc, c.dpart);
    hipLaunchKernelGGL(k_reduce_final<OpMax>, dim3(1), dim3(RED_FINAL_THREADS), 0, c.stream, (const double*)c.dpart, nb, 0.,
                       c.dscal + DS_VMAX);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// ---------------------------------------------------------------- initial velocity
// init_vel_p 'cou' (init.f90:191-210): v = v_lower + shear_rate*(k-0.5)*dz, k = 0..k1
template <class T>
__global__ void k_init_couette(Grid g, double v_lower, double shear_rate, T* __restrict__ v) {
    int i0 = blockIdx.x * BX + threadIdx.x, j0 = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i0 >= g.nx || j0 >= g.ny) return;
    double coorz = (k - 0.5) * g.dz;
    v[gidx(g, i0, j0, k)] = (T)(v_lower + shear_rate * coorz);
}

int launch_init_velocity(Ctx& c) {
    if (c.P.iniu != ICLS_INIU_COU) return ICLS_OK;
    double v_upper = c.P.wall_speed, v_lower = -c.P.wall_speed;
    double shear_rate = (v_upper - v_lower) / c.P.lz;
    dim3 blk(BX, BY, 1);
    dim3 grd = grid3(c.g, c.g.nz + 2, blk);
    if (c.fp32)
        hipLaunchKernelGGL(k_init_couette<float>, grd, blk, 0, c.stream, c.g, v_lower, shear_rate, as<float>(c.vnew));
    else
        hipLaunchKernelGGL(k_init_couette<double>, grd, blk, 0, c.stream, c.g, v_lower, shear_rate, c.vnew);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// fill a whole padded field (used instead of memset so graph capture stays safe)
template <class T>
__global__ void k_fill(long n, T* __restrict__ f, T v) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) f[t] = v;
}

int launch_fill(Ctx& c, double* f, double v, int fp32) {
    long n = c.nfield;
    int nb = (int)((n + 255) / 256);
    if (nb > 65536) nb = 65536;
    if (fp32) hipLaunchKernelGGL(k_fill<float>, dim3(nb), dim3(256), 0, c.stream, n, as<float>(f), (float)v);
    else hipLaunchKernelGGL(k_fill<double>, dim3(nb), dim3(256), 0, c.stream, n, f, v);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// element-type conversion of a whole padded field
template <class S, class D>
__global__ void k_convert(long n, const S* __restrict__ a, D* __restrict__ b) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) b[t] = (D)a[t];
}

int launch_convert(Ctx& c, const double* src, int src_fp32, double* dst, int dst_fp32) {
    long n = c.nfield;
    int nb = (int)((n + 255) / 256);
    if (nb > 65536) nb = 65536;
    if (src_fp32 && dst_fp32)
        hipLaunchKernelGGL((k_convert<float, float>), dim3(nb), dim3(256), 0, c.stream, n, as<float>(src), as<float>(dst));
    else if (src_fp32)
        hipLaunchKernelGGL((k_convert<float, double>), dim3(nb), dim3(256), 0, c.stream, n, as<float>(src), dst);
    else if (dst_fp32)
        hipLaunchKernelGGL((k_convert<double, float>), dim3(nb), dim3(256), 0, c.stream, n, src, as<float>(dst));
    else
        hipLaunchKernelGGL((k_convert<double, double>), dim3(nb), dim3(256), 0, c.stream, n, src, dst);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// copy of a whole padded field-class array
int launch_copy_field(Ctx& c, const double* a, double* b) { return launch_convert(c, a, c.fp32, b, c.fp32); }

}  // namespace icls
