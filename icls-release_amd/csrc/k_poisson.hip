// k_poisson.hip -- solver2d (solver.f90:68-196): real FFT along x, real FFT along y,
// Thomas elimination along z per Fourier mode, inverse FFTs.  On one GPU the
// 2DECOMP pencil transposes (solver.f90:82-83, 98, 113, 165, 179, 192-193) are
// identities: every pass works on the same i-fastest array.
//
// The transforms follow VRFFTF / VRFFTB semantics (libs/fft/fft.f:1073-1145,
// 374-450): half-complex packing R(1) = Re F0, R(2k) = Re Fk, R(2k+1) = Im Fk,
// R(N) = Re F(N/2), scaled by sqrt(1/N) both ways.  A length-N real line is
// transformed as a length-N/2 complex Stockham FFT in LDS plus the even/odd split.
//
// Precision.  FP64 mode: every pass is double and in place on p.  Mixed modes: the
// transforms run in float on the work array pf, the Thomas elimination runs in double
// on float storage, the last inverse transform writes double p, and one step of
// iterative refinement follows: r = rhs - L p in double (L is the discrete operator
// the transforms diagonalise), the same solve on r, p += the correction.
#include <type_traits>
#include <vector>

#include "icls_internal.hpp"

namespace icls {

template <class C> struct cx { C re, im; };
template <class C> __device__ __forceinline__ cx<C> cmul(cx<C> a, cx<C> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class C> __device__ __forceinline__ cx<C> cadd(cx<C> a, cx<C> b) { return {a.re + b.re, a.im + b.im}; }
template <class C> __device__ __forceinline__ cx<C> csub(cx<C> a, cx<C> b) { return {a.re - b.re, a.im - b.im}; }
template <class C> __device__ __forceinline__ cx<C> cconj(cx<C> a) { return {a.re, -a.im}; }

// out[q] = v, or with ACC out[q] = b0[q] + v: the refinement pass adds its correction
// to its float starting solution b0, which then takes the new solution rounded to float
// (the next step's starting solution)
template <class TO, class C, int ACC, class B>
__device__ __forceinline__ void put(TO* out, long q, C v, B* __restrict__ b0) {
    if (ACC) {
        const TO r = (TO)b0[q] + (TO)v;
        out[q] = r;
        b0[q] = (B)r;
    } else {
        out[q] = (TO)v;
    }
}

// Residual of the discrete Poisson problem the transforms and the Thomas solve
// invert exactly (periodic 3-point second differences in x and y, the Neumann
// tridiagonal rows of initsolver.f90:36-49 in z): rhs - L p at one cell, in double.
// From the cell's value pc, its x, y and z neighbours (zp unused at k = nz, zm at
// k = 1) and the right-hand side r.
__device__ __forceinline__ double residual_of(const Grid& g, int k, double r, double pc, double xp, double xm,
                                              double yp, double ym, double zp, double zm) {
    const double ax = g.dxi * g.dxi, ay = g.dyi * g.dyi, az = g.dzi * g.dzi;
    double lx = (xp - 2. * pc + xm) * ax;
    double ly = (yp - 2. * pc + ym) * ay;
    double lz;
    if (k == 1) lz = (zp - pc) * az;
    else if (k == g.nz) lz = (zm - pc) * az;
    else lz = (zp - 2. * pc + zm) * az;
    return r - (lx + ly + lz);
}
template <class T, class P>
__device__ __forceinline__ double residual_at(const Grid& g, const T* __restrict__ rhs, const P* __restrict__ p,
                                              int i0, int j0, int k) {
    const long q = gidx(g, i0, j0, k);
    const double zp = k < g.nz || k == 1 ? (double)p[q + g.sz] : 0.;
    const double zm = k > 1 ? (double)p[q - g.sz] : 0.;
    return residual_of(g, k, (double)rhs[q], (double)p[q], (double)p[gidx(g, wrap(i0 + 1, g.nx), j0, k)],
                       (double)p[gidx(g, wrap(i0 - 1, g.nx), j0, k)], (double)p[gidx(g, i0, wrap(j0 + 1, g.ny), k)],
                       (double)p[gidx(g, i0, wrap(j0 - 1, g.ny), k)], zp, zm);
}

// In-LDS Stockham radix-2 FFT of M points for one line; TPL threads per line.
// tw[t] = exp(+2 pi i t / N) for t < N/2 (N = 2M); sign = -1 forward, +1 inverse.
template <class C>
__device__ __forceinline__ void stockham(cx<C>* x, cx<C>* y, int M, int lane, int TPL, const cx<C>* tw, int sign) {
    for (int Ns = 1; Ns < M; Ns <<= 1) {
        for (int j = lane; j < M / 2; j += TPL) {
            int kk = j & (Ns - 1);
            cx<C> a = x[j], b = x[j + M / 2];
            cx<C> w = tw[kk * (M / Ns)];      // exp(+2 pi i kk / (2 Ns))
            if (sign < 0) w.im = -w.im;
            b = cmul(b, w);
            int d = ((j - kk) << 1) + kk;
            y[d] = cadd(a, b);
            y[d + Ns] = csub(a, b);
        }
        __syncthreads();
        cx<C>* t = x; x = y; y = t;
    }
}

// Runtime-sized transform (any power-of-two N).
// DIR 0: lines along x (contiguous), blocks of LPB consecutive j at one k.
// DIR 1: lines along y (stride nx), blocks of LPB consecutive i at one k.
// The input is C (the compute type), the output TO; ACC adds into the output.
template <int DIR, int INV, class C, class TO, int ACC, class B>
__global__ void k_rfft(Grid g, int N, int LPB, int TPL, const C* __restrict__ twg, const C* in, TO* out,
                       B* __restrict__ b0) {
    extern __shared__ unsigned char smem_raw[];
    cx<C>* smem = (cx<C>*)smem_raw;
    const int M = N / 2;
    cx<C>* bufa = smem;
    cx<C>* bufb = smem + (long)LPB * M;
    cx<C>* tw = smem + 2L * LPB * M;            // N/2 twiddles
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int nbl = (DIR == 0 ? g.ny : g.nx) / LPB;
    const int tile = xcd_tile(blockIdx.x, nbl * g.nz);
    const int k = tile / nbl + 1;
    const int line0 = (tile % nbl) * LPB;
    for (int t = tid; t < M; t += nthr) tw[t] = {twg[2 * t], twg[2 * t + 1]};
    const C scale = (C)sqrt(1. / (double)N);
    auto gq = [&](int ln, int n) -> long { return DIR == 0 ? gidx(g, n, line0 + ln, k) : gidx(g, line0 + ln, n, k); };
    auto split = [&](int e, int& ln, int& n) {
        if (DIR == 0) { ln = e / N; n = e - ln * N; } else { n = e / LPB; ln = e - n * LPB; }
    };
    if (!INV) {
        for (int e = tid; e < LPB * N; e += nthr) {
            int ln, n;
            split(e, ln, n);
            C v = in[gq(ln, n)];
            cx<C>* z = bufa + (long)ln * M + (n >> 1);
            if (n & 1) z->im = v; else z->re = v;
        }
    } else {
        // half-complex -> X_k, then Z_k = (X_k + conj X_{M-k}) + i w^k (X_k - conj X_{M-k})
        for (int e = tid; e < LPB * N; e += nthr) {
            int ln, n;
            split(e, ln, n);
            ((C*)(bufb + (long)ln * M))[n] = in[gq(ln, n)];
        }
        __syncthreads();
        for (int e = tid; e < LPB * M; e += nthr) {
            int ln = e / M, kk = e - ln * M;
            const C* r = (const C*)(bufb + (long)ln * M);
            cx<C> X, Y;   // X_k and X_{M-k}
            if (kk == 0) { X = {r[0], C(0.)}; Y = {r[N - 1], C(0.)}; }
            else { X = {r[2 * kk - 1], r[2 * kk]}; Y = {r[2 * (M - kk) - 1], r[2 * (M - kk)]}; }
            cx<C> cy = cconj(Y);
            cx<C> s = cadd(X, cy), dlt = csub(X, cy);
            cx<C> t = cmul(tw[kk], dlt);       // w^k (X_k - conj X_{M-k})
            cx<C> it = {-t.im, t.re};          // i * t
            bufa[(long)ln * M + kk] = cadd(s, it);
        }
    }
    __syncthreads();
    // ---- transform each line
    const int ln = tid / TPL, lane = tid - ln * TPL;
    if (ln < LPB) stockham(bufa + (long)ln * M, bufb + (long)ln * M, M, lane, TPL, tw, INV ? 1 : -1);
    else {
        // threads beyond LPB*TPL still take part in the barriers
        for (int Ns = 1; Ns < M; Ns <<= 1) __syncthreads();
    }
    // every line ends in the same buffer
    int nst = 0;
    for (int Ns = 1; Ns < M; Ns <<= 1) nst++;
    cx<C>* outb = (nst & 1) ? bufb : bufa;
    // ---- store
    if (!INV) {
        // X_k = (Z_k + conj Z_{M-k})/2 + w^-k (Z_k - conj Z_{M-k})/(2i), k = 0..M
        for (int e = tid; e < LPB * N; e += nthr) {
            int lnn, n;
            split(e, lnn, n);
            const cx<C>* Z = outb + (long)lnn * M;
            C v;
            if (n == 0) v = Z[0].re + Z[0].im;
            else if (n == N - 1) v = Z[0].re - Z[0].im;
            else {
                int kk = (n + 1) >> 1;
                cx<C> a = Z[kk], b = cconj(Z[M - kk]);
                cx<C> ev = {(a.re + b.re) * C(0.5), (a.im + b.im) * C(0.5)};
                cx<C> od = {(a.im - b.im) * C(0.5), -(a.re - b.re) * C(0.5)};   // (a - b)/(2i)
                cx<C> w = tw[kk];
                w.im = -w.im;
                cx<C> X = cadd(ev, cmul(w, od));
                v = (n & 1) ? X.re : X.im;
            }
            put<TO, C, ACC>(out, gq(lnn, n), scale * v, b0);
        }
    } else {
        for (int e = tid; e < LPB * N; e += nthr) {
            int lnn, n;
            split(e, lnn, n);
            cx<C> z = outb[(long)lnn * M + (n >> 1)];
            put<TO, C, ACC>(out, gq(lnn, n), scale * ((n & 1) ? z.im : z.re), b0);
        }
    }
}

// The same transform with the line length N and the lines per block LPB fixed at
// compile time: every index split is a shift and every stage loop unrolls.  With
// RES (forward x only) the input is not read from in but formed in double as the
// residual rhs - L b0 of the first pass's float solution b0, so the refinement pass
// starts without a residual array.  With ACC (inverse x only) the output is b0 plus
// the transform.
template <int N, int LPB>
struct RfftShape {
    static constexpr int M = N / 2;
    static constexpr int TPL = (M / 2) < 64 ? (M / 2) : 64;
    static constexpr int NTHR = LPB * TPL < 64 ? 64 : LPB * TPL;
};

template <int DIR, int INV, int N, int LPB, class C, class TO, int ACC, int RES, class B>
__global__ void __launch_bounds__((RfftShape<N, LPB>::NTHR))
k_rfft_n(Grid g, const C* __restrict__ twg, const C* in, TO* out, B* __restrict__ b0) {
    static_assert(!RES || (DIR == 0 && !INV), "the residual feeds the forward x transform");
    constexpr int M = N / 2;
    constexpr int TPL = RfftShape<N, LPB>::TPL;
    constexpr int NTHR = RfftShape<N, LPB>::NTHR;
    __shared__ cx<C> bufa[LPB * M];
    __shared__ cx<C> bufb[LPB * M];
    __shared__ cx<C> tw[M];
    const int tid = threadIdx.x;
    const int nbl = (DIR == 0 ? g.ny : g.nx) / LPB;           // blocks per z layer
    const int tile = xcd_tile(blockIdx.x, nbl * g.nz);
    const int k = tile / nbl + 1;
    const int line0 = (tile % nbl) * LPB;
    for (int t = tid; t < M; t += NTHR) tw[t] = {twg[2 * t], twg[2 * t + 1]};
    const C scale = (C)sqrt(1. / (double)N);
    auto gq = [&](int ln, int n) -> long { return DIR == 0 ? gidx(g, n, line0 + ln, k) : gidx(g, line0 + ln, n, k); };
    auto split = [&](int e, int& ln, int& n) {
        if (DIR == 0) { ln = e / N; n = e % N; } else { n = e / LPB; ln = e % LPB; }
    };
    if (!INV) {
#pragma unroll 4
        for (int e = tid; e < LPB * N; e += NTHR) {
            int ln, n;
            split(e, ln, n);
            C v;
            if (RES) v = (C)residual_at(g, in, b0, n, line0 + ln, k);
            else v = in[gq(ln, n)];
            cx<C>* z = bufa + ln * M + (n >> 1);
            if (n & 1) z->im = v; else z->re = v;
        }
    } else {
#pragma unroll 4
        for (int e = tid; e < LPB * N; e += NTHR) {
            int ln, n;
            split(e, ln, n);
            ((C*)(bufb + ln * M))[n] = in[gq(ln, n)];
        }
        __syncthreads();
        for (int e = tid; e < LPB * M; e += NTHR) {
            int ln = e / M, kk = e % M;
            const C* r = (const C*)(bufb + ln * M);
            cx<C> X, Y;
            if (kk == 0) { X = {r[0], C(0.)}; Y = {r[N - 1], C(0.)}; }
            else { X = {r[2 * kk - 1], r[2 * kk]}; Y = {r[2 * (M - kk) - 1], r[2 * (M - kk)]}; }
            cx<C> cy = cconj(Y);
            cx<C> s = cadd(X, cy), dlt = csub(X, cy);
            cx<C> t = cmul(tw[kk], dlt);
            cx<C> it = {-t.im, t.re};
            bufa[ln * M + kk] = cadd(s, it);
        }
    }
    __syncthreads();
    const int ln = tid / TPL, lane = tid % TPL;
    cx<C>* x = bufa + (ln < LPB ? ln : 0) * M;
    cx<C>* y = bufb + (ln < LPB ? ln : 0) * M;
#pragma unroll
    for (int Ns = 1; Ns < M; Ns <<= 1) {
        if (ln < LPB) {
#pragma unroll
            for (int j = lane; j < M / 2; j += TPL) {
                int kk = j & (Ns - 1);
                cx<C> a = x[j], b = x[j + M / 2];
                cx<C> w = tw[kk * (M / Ns)];
                if (!INV) w.im = -w.im;
                b = cmul(b, w);
                int d = ((j - kk) << 1) + kk;
                y[d] = cadd(a, b);
                y[d + Ns] = csub(a, b);
            }
        }
        __syncthreads();
        cx<C>* t = x; x = y; y = t;
    }
    constexpr int NST = (M >= 2) ? __builtin_ctz(M) : 0;
    cx<C>* outb = (NST & 1) ? bufb : bufa;
    if (!INV) {
#pragma unroll 4
        for (int e = tid; e < LPB * N; e += NTHR) {
            int lnn, n;
            split(e, lnn, n);
            const cx<C>* Z = outb + lnn * M;
            C v;
            if (n == 0) v = Z[0].re + Z[0].im;
            else if (n == N - 1) v = Z[0].re - Z[0].im;
            else {
                int kk = (n + 1) >> 1;
                cx<C> a = Z[kk], b = cconj(Z[M - kk]);
                cx<C> ev = {(a.re + b.re) * C(0.5), (a.im + b.im) * C(0.5)};
                cx<C> od = {(a.im - b.im) * C(0.5), -(a.re - b.re) * C(0.5)};
                cx<C> w = tw[kk];
                w.im = -w.im;
                cx<C> X = cadd(ev, cmul(w, od));
                v = (n & 1) ? X.re : X.im;
            }
            put<TO, C, ACC>(out, gq(lnn, n), scale * v, b0);
        }
    } else {
#pragma unroll 4
        for (int e = tid; e < LPB * N; e += NTHR) {
            int lnn, n;
            split(e, lnn, n);
            cx<C> z = outb[lnn * M + (n >> 1)];
            put<TO, C, ACC>(out, gq(lnn, n), scale * ((n & 1) ? z.im : z.re), b0);
        }
    }
}

// ---------------------------------------------------------------- register transform
// Float lines of length N = 256: the 128-point complex FFT of z_m = x_2m + i x_2m+1
// with its 128 points spread over 8 threads ("slots"), point m = s + 8 r in register
// r of slot s.  Each thread does a 16-point DFT over r in registers and applies the
// twiddles W128^(s k1); one LDS exchange gives every thread two k1 values from all 8
// slots for the 8-point DFTs (X_k1+16k2); a second exchange pairs X_k with X_128-k for
// the even/odd split (or, inverse, the split comes first, from global memory).
// Along y a block is 8 waves over 64 consecutive i: the lane is the line and the wave
// the slot, so every global access is a 256-byte row segment.  Along x a block is 4
// waves over 32 lines (consecutive j): 8 lanes per line, slot = lane % 8.
namespace r16 {
constexpr int N = 256, M = 128;

template <class T>
__device__ __forceinline__ cx<T> w16(int j, int sg) {
    // cos and sin of pi/8 and pi/4, in the element type
    constexpr T C1 = std::is_same<T, float>::value ? T(0.92387953251128674f) : T(0.92387953251128674);
    constexpr T S1 = std::is_same<T, float>::value ? T(0.38268343236508978f) : T(0.38268343236508978);
    constexpr T C2 = std::is_same<T, float>::value ? T(0.70710678118654752f) : T(0.70710678118654752);
    const T co[16] = {T(1.), C1, C2, S1, T(0.), -S1, -C2, -C1, T(-1.), -C1, -C2, -S1, T(0.), S1, C2, C1};
    const T si[16] = {T(0.), S1, C2, C1, T(1.), C1, C2, S1, T(0.), -S1, -C2, -C1, T(-1.), -C1, -C2, -S1};
    j &= 15;
    return {co[j], sg > 0 ? si[j] : -si[j]};
}

// 4-point DFT in place, exp(SG 2 pi i / 4) = SG i
template <int SG, class T>
__device__ __forceinline__ void dft4(cx<T>& x0, cx<T>& x1, cx<T>& x2, cx<T>& x3) {
    const cx<T> t0 = cadd(x0, x2), t1 = csub(x0, x2), t2 = cadd(x1, x3), t3 = csub(x1, x3);
    const cx<T> jt3 = SG > 0 ? cx<T>{-t3.im, t3.re} : cx<T>{t3.im, -t3.re};
    x0 = cadd(t0, t2);
    x1 = cadd(t1, jt3);
    x2 = csub(t0, t2);
    x3 = csub(t1, jt3);
}

// 16-point DFT of v[0..15] in place (4 x 4); X_k ends at v[4 (k % 4) + k / 4]
template <int SG, class T>
__device__ __forceinline__ void dft16(cx<T>* v) {
#pragma unroll
    for (int q = 0; q < 4; q++) dft4<SG>(v[q], v[q + 4], v[q + 8], v[q + 12]);
#pragma unroll
    for (int q = 1; q < 4; q++)
#pragma unroll
        for (int ka = 1; ka < 4; ka++) v[q + 4 * ka] = cmul(v[q + 4 * ka], w16<T>(q * ka, SG));
#pragma unroll
    for (int ka = 0; ka < 4; ka++) dft4<SG>(v[4 * ka], v[4 * ka + 1], v[4 * ka + 2], v[4 * ka + 3]);
}

// 8-point DFT of u[0..7] in place (4 x 2); X_k ends at u[2 (k % 4) + k / 4]
template <int SG, class T>
__device__ __forceinline__ void dft8(cx<T>* u) {
    dft4<SG>(u[0], u[2], u[4], u[6]);
    dft4<SG>(u[1], u[3], u[5], u[7]);
    u[3] = cmul(u[3], w16<T>(2, SG));
    u[5] = cmul(u[5], w16<T>(4, SG));
    u[7] = cmul(u[7], w16<T>(6, SG));
#pragma unroll
    for (int ka = 0; ka < 4; ka++) {
        const cx<T> a = u[2 * ka], b = u[2 * ka + 1];
        u[2 * ka] = cadd(a, b);
        u[2 * ka + 1] = csub(a, b);
    }
}

// exp(+2 pi i t / N) for 0 <= t < N from the half table
template <class T>
__device__ __forceinline__ cx<T> twf(const cx<T>* tw, int t) {
    return t < M ? tw[t] : cx<T>{-tw[t - M].re, -tw[t - M].im};
}

// Z_k of the inverse split from X_k and X_(M-k) of the half-complex line
template <class T>
__device__ __forceinline__ cx<T> inv_split(const cx<T>* tw, int kk, cx<T> X, cx<T> Y) {
    const cx<T> cy = cconj(Y);
    const cx<T> t = cmul(tw[kk], csub(X, cy));
    return cadd(cadd(X, cy), cx<T>{-t.im, t.re});
}

// F_k of the forward split from X_k and X_(M-k) (k >= 1)
template <class T>
__device__ __forceinline__ cx<T> fwd_split(const cx<T>* tw, int kk, cx<T> a, cx<T> y) {
    const cx<T> b = cconj(y);
    const cx<T> ev = {(a.re + b.re) * T(0.5), (a.im + b.im) * T(0.5)};
    const cx<T> od = {(a.im - b.im) * T(0.5), -(a.re - b.re) * T(0.5)};
    return cadd(ev, cmul(cconj(tw[kk]), od));
}

// the complex FFT from "v[r] = z_(s + 8 r)" to "u[h][2 (k2 % 4) + k2 / 4] = X_(2s+h+16k2)";
// xch(k1, t) addresses the exchange slot of (k1, slot t) of this thread's line
template <int SG, class T, class XCH>
__device__ __forceinline__ void cfft128(cx<T>* v, cx<T> (*u)[8], int s, const cx<T>* tw, XCH xch) {
    dft16<SG>(v);
#pragma unroll
    for (int k1 = 0; k1 < 16; k1++) {
        cx<T> val = v[4 * (k1 % 4) + k1 / 4];
        if (k1) {
            cx<T> w = twf(tw, 2 * s * k1);
            if (SG < 0) w.im = -w.im;
            val = cmul(val, w);
        }
        *xch(k1, s) = val;
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; h++) {
#pragma unroll
        for (int t = 0; t < 8; t++) u[h][t] = *xch(2 * s + h, t);
        dft8<SG>(u[h]);
    }
    __syncthreads();   // the exchange buffer is free again
}
}  // namespace r16

template <class TO, int ACC, class B>
__device__ __forceinline__ void put2(TO* out, long q, float a, float b, B* __restrict__ b0) {
    put<TO, float, ACC>(out, q, a, b0);
    put<TO, float, ACC>(out, q + 1, b, b0);
}

// along y: 512 threads, 64 lines (i) x 8 slots; LDS 64 KB (float) or 128 KB (double)
// exchange + twiddles
template <int INV, class C, class TO, int ACC>
__global__ void __launch_bounds__(512) k_rfft256_y(Grid g, const C* __restrict__ twg, const C* in, TO* out) {
    using namespace r16;
    static_assert(!ACC, "only the x transform adds into its output");
    __shared__ cx<C> xb[M * 64];
    __shared__ cx<C> tw[M];
    const int lane = threadIdx.x & 63, s = threadIdx.x >> 6;
    const int nbl = g.nx / 64;
    const int tile = xcd_tile(blockIdx.x, nbl * g.nz);
    const int k = tile / nbl + 1, i = (tile % nbl) * 64 + lane;
    for (int t = threadIdx.x; t < M; t += 512) tw[t] = {twg[2 * t], twg[2 * t + 1]};
    const C scale = C(0.0625);   // sqrt(1/256)
    auto row = [&](int n) -> long { return gidx(g, i, n, k); };
    cx<C> v[16];
    if (!INV) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int m = s + 8 * r;
            v[r] = {in[row(2 * m)], in[row(2 * m + 1)]};
        }
        __syncthreads();
    } else {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int kk = s + 8 * r;
            cx<C> X, Y;
            if (kk == 0) { X = {in[row(0)], C(0.)}; Y = {in[row(N - 1)], C(0.)}; }
            else {
                X = {in[row(2 * kk - 1)], in[row(2 * kk)]};
                Y = {in[row(2 * (M - kk) - 1)], in[row(2 * (M - kk))]};
            }
            v[r] = inv_split(tw, kk, X, Y);
        }
    }
    cx<C> u[2][8];
    auto xch = [&](int k1, int t) { return &xb[(k1 * 8 + t) * 64 + lane]; };
    if (!INV) cfft128<-1>(v, u, s, tw, xch);
    else cfft128<1>(v, u, s, tw, xch);
    if (INV) {
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int k2 = 0; k2 < 8; k2++) {
                const int m = 2 * s + h + 16 * k2;
                const cx<C> z = u[h][2 * (k2 % 4) + k2 / 4];
                put<TO, C, ACC>(out, row(2 * m), z.re * scale, (C*)nullptr);
                put<TO, C, ACC>(out, row(2 * m + 1), z.im * scale, (C*)nullptr);
            }
        return;
    }
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int k2 = 0; k2 < 8; k2++) xb[((2 * s + h) + 16 * k2) * 64 + lane] = u[h][2 * (k2 % 4) + k2 / 4];
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int k2 = 0; k2 < 8; k2++) {
            const int kk = 2 * s + h + 16 * k2;
            const cx<C> a = u[h][2 * (k2 % 4) + k2 / 4];
            if (kk == 0) {
                put<TO, C, ACC>(out, row(0), (a.re + a.im) * scale, (C*)nullptr);
                put<TO, C, ACC>(out, row(N - 1), (a.re - a.im) * scale, (C*)nullptr);
            } else {
                const cx<C> F = fwd_split(tw, kk, a, xb[(M - kk) * 64 + lane]);
                put<TO, C, ACC>(out, row(2 * kk - 1), F.re * scale, (C*)nullptr);
                put<TO, C, ACC>(out, row(2 * kk), F.im * scale, (C*)nullptr);
            }
        }
}

struct cx4 {
    float v[4];
};

// along x: 256 threads, 32 lines (j) x 8 slots; with RES (forward) the input is
// rhs - L b0 formed on the fly.  C = double: the FP64 transforms in place on p
// (neither RES nor ACC).
template <int INV, class C, class TO, int ACC, int RES, class B>
__global__ void __launch_bounds__(256) k_rfft256_x(Grid g, const C* __restrict__ twg, const C* in, TO* out,
                                                  B* __restrict__ b0) {
    using namespace r16;
    static_assert(!RES || !INV, "the residual feeds the forward transform");
    static_assert(INV || (!ACC && sizeof(TO) == sizeof(C)), "the forward transform writes the work array");
    static_assert(std::is_same<C, float>::value || (!RES && !ACC && std::is_same<TO, C>::value),
                  "the FP64 transforms neither form the residual nor accumulate");
    constexpr int LS = 16 * 8 + 8;    // exchange stride of a line (complex)
    __shared__ cx<C> xb[32 * LS];
    __shared__ cx<C> tw[M];
    const int lane = threadIdx.x & 63, s = lane & 7, line = (threadIdx.x >> 6) * 8 + (lane >> 3);
    const int nbl = g.ny / 32;
    const int tile = xcd_tile(blockIdx.x, nbl * g.nz);
    const int k = tile / nbl + 1, j = (tile % nbl) * 32 + line;
    for (int t = threadIdx.x; t < M; t += 256) tw[t] = {twg[2 * t], twg[2 * t + 1]};
    const C scale = C(0.0625);
    const long q0 = gidx(g, 0, j, k);
    cx<C>* xl = xb + line * LS;
    cx<C> v[16];
    if constexpr (RES) {
        // the residual line goes through LDS so that its stencils are not all in flight;
        // the neighbour lines are addressed from their starts.  A thread forms four
        // consecutive cells from 16-byte loads of the line and of its four neighbour
        // lines (and two loads for the x neighbours of the four), so a wavefront's loads
        // cover whole 128-byte rows.
        static_assert(std::is_same<B, float>::value, "the starting solution is float");
        float* fl = reinterpret_cast<float*>(xl);
        const float* pc = b0 + q0;
        const float* pyp = b0 + gidx(g, 0, wrap(j + 1, g.ny), k);
        const float* pym = b0 + gidx(g, 0, wrap(j - 1, g.ny), k);
        const bool zup = k < g.nz || k == 1, zdn = k > 1;
        const float* pzp = pc + (zup ? g.sz : 0);
        const float* pzm = pc - (zdn ? g.sz : 0);
        const float* rl = in + q0;
        auto ld4 = [](const float* a) {
            const float4 t = *reinterpret_cast<const float4*>(a);
            return cx4{t.x, t.y, t.z, t.w};
        };
#pragma unroll 2
        for (int it = 0; it < N / 32; it++) {
            const int x = 4 * s + 32 * it;
            const cx4 r4 = ld4(rl + x), c4 = ld4(pc + x), yp4 = ld4(pyp + x), ym4 = ld4(pym + x);
            const cx4 zp4 = ld4(pzp + x), zm4 = ld4(pzm + x);
            const float cl = pc[(x - 1) & (N - 1)], cr = pc[(x + 4) & (N - 1)];
            const float cc[6] = {cl, c4.v[0], c4.v[1], c4.v[2], c4.v[3], cr};
            float o4[4];
#pragma unroll
            for (int t = 0; t < 4; t++)
                o4[t] = (float)residual_of(g, k, (double)r4.v[t], (double)cc[t + 1], (double)cc[t + 2], (double)cc[t],
                                           (double)yp4.v[t], (double)ym4.v[t], (double)zp4.v[t], (double)zm4.v[t]);
            *reinterpret_cast<float4*>(fl + x) = make_float4(o4[0], o4[1], o4[2], o4[3]);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; r++) v[r] = xl[s + 8 * r];
        __syncthreads();
    } else if (!INV) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            if constexpr (std::is_same<C, float>::value) {
                const float2 a = *reinterpret_cast<const float2*>(in + q0 + 2 * (s + 8 * r));
                v[r] = {a.x, a.y};
            } else {
                const double2 a = *reinterpret_cast<const double2*>(in + q0 + 2 * (s + 8 * r));
                v[r] = {a.x, a.y};
            }
        }
        __syncthreads();
    } else {
        // the half-complex line goes through LDS, loaded with 16-byte vectors: the split
        // reads it at scattered positions
        C* fl = reinterpret_cast<C*>(xl);
        constexpr int VW = 16 / sizeof(C);   // reals per vector
#pragma unroll
        for (int it = 0; it < N / (8 * VW); it++) {
            const int x = VW * s + 8 * VW * it;
            if constexpr (VW == 4) *reinterpret_cast<float4*>(fl + x) = *reinterpret_cast<const float4*>(in + q0 + x);
            else *reinterpret_cast<double2*>(fl + x) = *reinterpret_cast<const double2*>(in + q0 + x);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int kk = s + 8 * r;
            cx<C> X, Y;
            if (kk == 0) { X = {fl[0], C(0.)}; Y = {fl[N - 1], C(0.)}; }
            else {
                X = {fl[2 * kk - 1], fl[2 * kk]};
                Y = {fl[2 * (M - kk) - 1], fl[2 * (M - kk)]};
            }
            v[r] = inv_split(tw, kk, X, Y);
        }
        __syncthreads();   // the line buffer is the exchange buffer of the transform
    }
    cx<C> u[2][8];
    auto xch = [&](int k1, int t) { return &xl[k1 * 8 + t]; };
    if (!INV) cfft128<-1>(v, u, s, tw, xch);
    else cfft128<1>(v, u, s, tw, xch);
    if (INV) {
        // x = 4 s + 32 k2 .. +3 holds z_(2s+16k2) and z_(2s+1+16k2)
#pragma unroll
        for (int k2 = 0; k2 < 8; k2++) {
            const cx<C> z0 = u[0][2 * (k2 % 4) + k2 / 4], z1 = u[1][2 * (k2 % 4) + k2 / 4];
            const long q = q0 + 4 * s + 32 * k2;
            if constexpr (std::is_same<C, double>::value) {
                *reinterpret_cast<double2*>(out + q) = make_double2(z0.re * scale, z0.im * scale);
                *reinterpret_cast<double2*>(out + q + 2) = make_double2(z1.re * scale, z1.im * scale);
            } else if constexpr (sizeof(TO) == 4 && !ACC) {
                *reinterpret_cast<float4*>(out + q) = make_float4(z0.re * scale, z0.im * scale, z1.re * scale, z1.im * scale);
            } else if constexpr (sizeof(TO) == 8 && ACC) {
                // q is a multiple of 4: the base and the output move as 16-byte vectors
                static_assert(std::is_same<B, float>::value, "the starting solution is float");
                const float4 t = *reinterpret_cast<const float4*>(b0 + q);
                const double r0 = (double)t.x + (double)(z0.re * scale), r1 = (double)t.y + (double)(z0.im * scale);
                const double r2 = (double)t.z + (double)(z1.re * scale), r3 = (double)t.w + (double)(z1.im * scale);
                *reinterpret_cast<double2*>(out + q) = make_double2(r0, r1);
                *reinterpret_cast<double2*>(out + q + 2) = make_double2(r2, r3);
                *reinterpret_cast<float4*>(b0 + q) = make_float4((float)r0, (float)r1, (float)r2, (float)r3);
            } else {
                put2<TO, ACC>(out, q, z0.re * scale, z0.im * scale, b0);
                put2<TO, ACC>(out, q + 2, z1.re * scale, z1.im * scale, b0);
            }
        }
        return;
    }
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int k2 = 0; k2 < 8; k2++) xl[(2 * s + h) + 16 * k2] = u[h][2 * (k2 % 4) + k2 / 4];
    __syncthreads();
    C o[2][8][2];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int k2 = 0; k2 < 8; k2++) {
            const int kk = 2 * s + h + 16 * k2;
            const cx<C> a = u[h][2 * (k2 % 4) + k2 / 4];
            if (kk == 0) {
                o[h][k2][0] = (a.re + a.im) * scale;   // x = 0
                o[h][k2][1] = (a.re - a.im) * scale;   // x = N - 1
            } else {
                const cx<C> F = fwd_split(tw, kk, a, xl[M - kk]);
                o[h][k2][0] = F.re * scale;            // x = 2 kk - 1
                o[h][k2][1] = F.im * scale;            // x = 2 kk
            }
        }
    __syncthreads();
    // stage the line in LDS (reals) for aligned, coalesced stores
    C* fl = reinterpret_cast<C*>(xl);
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int k2 = 0; k2 < 8; k2++) {
            const int kk = 2 * s + h + 16 * k2;
            if (kk == 0) { fl[0] = o[h][k2][0]; fl[N - 1] = o[h][k2][1]; }
            else { fl[2 * kk - 1] = o[h][k2][0]; fl[2 * kk] = o[h][k2][1]; }
        }
    __syncthreads();
    if constexpr (std::is_same<C, float>::value) {
        // 32 lines x 64 float4: thread t stores float4 t % 64 of lines t / 64 + 4 it
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int ln = (threadIdx.x >> 6) + 4 * it, c4 = threadIdx.x & 63;
            const float* src = reinterpret_cast<const float*>(xb + ln * LS) + 4 * c4;
            const long q = gidx(g, 4 * c4, (tile % nbl) * 32 + ln, k);
            *reinterpret_cast<float4*>(out + q) = make_float4(src[0], src[1], src[2], src[3]);
        }
    } else {
        // 32 lines x 128 double2: thread t stores double2 t % 128 of lines t / 128 + 2 it
#pragma unroll
        for (int it = 0; it < 16; it++) {
            const int ln = (threadIdx.x >> 7) + 2 * it, c2 = threadIdx.x & 127;
            const double* src = reinterpret_cast<const double*>(xb + ln * LS) + 2 * c2;
            const long q = gidx(g, 2 * c2, (tile % nbl) * 32 + ln, k);
            *reinterpret_cast<double2*>(out + q) = make_double2(src[0], src[1]);
        }
    }
}

// ---------------------------------------------------------------- register transform, 512
// Float lines of length N = 512: the 256-point complex FFT of z_m = x_2m + i x_2m+1
// over 16 slots, point m = s + 16 r in register r of slot s.  A 16-point DFT over r
// in registers and the twiddles W256^(s k1); one LDS exchange after which slot s holds
// the 16 values of k1 = s from all slots; a second 16-point DFT gives X_(s + 16 k2).
// A further exchange pairs X_k with X_256-k for the even/odd split (inverse: the split
// comes first, from global memory).  Along y a block is 16 waves over 64 consecutive i
// (lane = line, wave = slot); along x 4 waves over 16 lines, 16 lanes per line.
namespace r512 {
constexpr int N = 512, M = 256;

// exp(+2 pi i t / N) for 0 <= t < N from the half table
template <class T>
__device__ __forceinline__ cx<T> twf(const cx<T>* tw, int t) {
    return t < M ? tw[t] : cx<T>{-tw[t - M].re, -tw[t - M].im};
}

// the complex FFT from "v[r] = z_(s + 16 r)" to "v[4 (k2 % 4) + k2 / 4] = X_(s + 16 k2)";
// xch(k1, t) addresses the exchange slot of (k1, slot t) of this thread's line
template <int SG, class T, class XCH>
__device__ __forceinline__ void cfft256(cx<T>* v, int s, const cx<T>* tw, XCH xch) {
    r16::dft16<SG>(v);
#pragma unroll
    for (int k1 = 0; k1 < 16; k1++) {
        cx<T> val = v[4 * (k1 % 4) + k1 / 4];
        if (k1) {
            cx<T> w = twf(tw, 2 * s * k1);
            if (SG < 0) w.im = -w.im;
            val = cmul(val, w);
        }
        *xch(k1, s) = val;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 16; t++) v[t] = *xch(s, t);
    r16::dft16<SG>(v);
    __syncthreads();   // the exchange buffer is free again
}
}  // namespace r512

// along y: 1024 threads, 64 lines (i) x 16 slots; LDS 128 KB exchange + twiddles
template <int INV, class TO, int ACC>
__global__ void __launch_bounds__(1024) k_rfft512_y(Grid g, const float* __restrict__ twg, const float* in, TO* out) {
    using namespace r512;
    static_assert(!ACC, "only the x transform adds into its output");
    __shared__ cx<float> xb[M * 64];
    __shared__ cx<float> tw[M];
    const int lane = threadIdx.x & 63, s = threadIdx.x >> 6;
    const int nbl = g.nx / 64;
    const int tile = xcd_tile(blockIdx.x, nbl * g.nz);
    const int k = tile / nbl + 1, i = (tile % nbl) * 64 + lane;
    for (int t = threadIdx.x; t < M; t += 1024) tw[t] = {twg[2 * t], twg[2 * t + 1]};
    const float scale = (float)sqrt(1. / (double)N);
    auto row = [&](int n) -> long { return gidx(g, i, n, k); };
    cx<float> v[16];
    if (!INV) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int m = s + 16 * r;
            v[r] = {in[row(2 * m)], in[row(2 * m + 1)]};
        }
        __syncthreads();
    } else {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int kk = s + 16 * r;
            cx<float> X, Y;
            if (kk == 0) { X = {in[row(0)], 0.f}; Y = {in[row(N - 1)], 0.f}; }
            else {
                X = {in[row(2 * kk - 1)], in[row(2 * kk)]};
                Y = {in[row(2 * (M - kk) - 1)], in[row(2 * (M - kk))]};
            }
            v[r] = r16::inv_split(tw, kk, X, Y);
        }
    }
    auto xch = [&](int k1, int t) { return &xb[(k1 * 16 + t) * 64 + lane]; };
    if (!INV) cfft256<-1>(v, s, tw, xch);
    else cfft256<1>(v, s, tw, xch);
    if (INV) {
#pragma unroll
        for (int k2 = 0; k2 < 16; k2++) {
            const int m = s + 16 * k2;
            const cx<float> z = v[4 * (k2 % 4) + k2 / 4];
            put<TO, float, ACC>(out, row(2 * m), z.re * scale, (float*)nullptr);
            put<TO, float, ACC>(out, row(2 * m + 1), z.im * scale, (float*)nullptr);
        }
        return;
    }
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) xb[(s + 16 * k2) * 64 + lane] = v[4 * (k2 % 4) + k2 / 4];
    __syncthreads();
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) {
        const int kk = s + 16 * k2;
        const cx<float> a = v[4 * (k2 % 4) + k2 / 4];
        if (kk == 0) {
            put<TO, float, ACC>(out, row(0), (a.re + a.im) * scale, (float*)nullptr);
            put<TO, float, ACC>(out, row(N - 1), (a.re - a.im) * scale, (float*)nullptr);
        } else {
            const cx<float> F = r16::fwd_split(tw, kk, a, xb[(M - kk) * 64 + lane]);
            put<TO, float, ACC>(out, row(2 * kk - 1), F.re * scale, (float*)nullptr);
            put<TO, float, ACC>(out, row(2 * kk), F.im * scale, (float*)nullptr);
        }
    }
}

// along x: 256 threads, 16 lines (j) x 16 slots; with RES (forward) the input is
// rhs - L b0 formed on the fly, as in the 256-point kernel.  C = double: the FP64
// transforms in place on p (neither RES nor ACC).
template <int INV, class C, class TO, int ACC, int RES, class B>
__global__ void __launch_bounds__(256) k_rfft512_x(Grid g, const C* __restrict__ twg, const C* in, TO* out,
                                                  B* __restrict__ b0) {
    using namespace r512;
    static_assert(!RES || !INV, "the residual feeds the forward transform");
    static_assert(INV || (!ACC && sizeof(TO) == sizeof(C)), "the forward transform writes the work array");
    static_assert(std::is_same<C, float>::value || (!RES && !ACC && std::is_same<TO, C>::value),
                  "the FP64 transforms neither form the residual nor accumulate");
    constexpr int LS = 16 * 17 + 4;   // line buffer (complex): the exchange at stride 17, padded
    __shared__ cx<C> xb[16 * LS];
    __shared__ cx<C> tw[M];
    const int lane = threadIdx.x & 63, s = lane & 15, line = (threadIdx.x >> 6) * 4 + (lane >> 4);
    const int nbl = g.ny / 16;
    const int tile = xcd_tile(blockIdx.x, nbl * g.nz);
    const int k = tile / nbl + 1, j = (tile % nbl) * 16 + line;
    for (int t = threadIdx.x; t < M; t += 256) tw[t] = {twg[2 * t], twg[2 * t + 1]};
    const C scale = (C)sqrt(1. / (double)N);
    const long q0 = gidx(g, 0, j, k);
    cx<C>* xl = xb + line * LS;
    cx<C> v[16];
    if constexpr (RES) {
        // four consecutive cells per thread from 16-byte loads of the line and of its
        // four neighbour lines, staged through LDS
        static_assert(std::is_same<B, float>::value, "the starting solution is float");
        float* fl = reinterpret_cast<float*>(xl);
        const float* pc = b0 + q0;
        const float* pyp = b0 + gidx(g, 0, wrap(j + 1, g.ny), k);
        const float* pym = b0 + gidx(g, 0, wrap(j - 1, g.ny), k);
        const bool zup = k < g.nz || k == 1, zdn = k > 1;
        const float* pzp = pc + (zup ? g.sz : 0);
        const float* pzm = pc - (zdn ? g.sz : 0);
        const float* rl = in + q0;
        auto ld4 = [](const float* a) {
            const float4 t = *reinterpret_cast<const float4*>(a);
            return cx4{t.x, t.y, t.z, t.w};
        };
#pragma unroll 2
        for (int it = 0; it < N / 64; it++) {
            const int x = 4 * s + 64 * it;
            const cx4 r4 = ld4(rl + x), c4 = ld4(pc + x), yp4 = ld4(pyp + x), ym4 = ld4(pym + x);
            const cx4 zp4 = ld4(pzp + x), zm4 = ld4(pzm + x);
            const float cl = pc[(x - 1) & (N - 1)], cr = pc[(x + 4) & (N - 1)];
            const float cc[6] = {cl, c4.v[0], c4.v[1], c4.v[2], c4.v[3], cr};
            float o4[4];
#pragma unroll
            for (int t = 0; t < 4; t++)
                o4[t] = (float)residual_of(g, k, (double)r4.v[t], (double)cc[t + 1], (double)cc[t + 2], (double)cc[t],
                                           (double)yp4.v[t], (double)ym4.v[t], (double)zp4.v[t], (double)zm4.v[t]);
            *reinterpret_cast<float4*>(fl + x) = make_float4(o4[0], o4[1], o4[2], o4[3]);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; r++) v[r] = xl[s + 16 * r];
        __syncthreads();
    } else if (!INV) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            if constexpr (std::is_same<C, float>::value) {
                const float2 a = *reinterpret_cast<const float2*>(in + q0 + 2 * (s + 16 * r));
                v[r] = {a.x, a.y};
            } else {
                const double2 a = *reinterpret_cast<const double2*>(in + q0 + 2 * (s + 16 * r));
                v[r] = {a.x, a.y};
            }
        }
        __syncthreads();
    } else {
        // the half-complex line through LDS, loaded with 16-byte vectors
        C* fl = reinterpret_cast<C*>(xl);
        constexpr int VW = 16 / sizeof(C);   // reals per vector
#pragma unroll
        for (int it = 0; it < N / (16 * VW); it++) {
            const int x = VW * s + 16 * VW * it;
            if constexpr (VW == 4) *reinterpret_cast<float4*>(fl + x) = *reinterpret_cast<const float4*>(in + q0 + x);
            else *reinterpret_cast<double2*>(fl + x) = *reinterpret_cast<const double2*>(in + q0 + x);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int kk = s + 16 * r;
            cx<C> X, Y;
            if (kk == 0) { X = {fl[0], C(0.)}; Y = {fl[N - 1], C(0.)}; }
            else {
                X = {fl[2 * kk - 1], fl[2 * kk]};
                Y = {fl[2 * (M - kk) - 1], fl[2 * (M - kk)]};
            }
            v[r] = r16::inv_split(tw, kk, X, Y);
        }
        __syncthreads();   // the line buffer is the exchange buffer of the transform
    }
    auto xch = [&](int k1, int t) { return &xl[k1 * 17 + t]; };
    if (!INV) cfft256<-1>(v, s, tw, xch);
    else cfft256<1>(v, s, tw, xch);
    if (INV) {
        // slot s holds z_m for m = s + 16 k2: x = 2 m and 2 m + 1
#pragma unroll
        for (int k2 = 0; k2 < 16; k2++) {
            const cx<C> z = v[4 * (k2 % 4) + k2 / 4];
            const long q = q0 + 2 * (s + 16 * k2);
            if constexpr (std::is_same<C, double>::value) {
                *reinterpret_cast<double2*>(out + q) = make_double2(z.re * scale, z.im * scale);
            } else if constexpr (sizeof(TO) == 4 && !ACC) {
                *reinterpret_cast<float2*>(out + q) = make_float2(z.re * scale, z.im * scale);
            } else if constexpr (sizeof(TO) == 8 && ACC) {
                static_assert(std::is_same<B, float>::value, "the starting solution is float");
                const float2 t = *reinterpret_cast<const float2*>(b0 + q);
                const double r0 = (double)t.x + (double)(z.re * scale), r1 = (double)t.y + (double)(z.im * scale);
                *reinterpret_cast<double2*>(out + q) = make_double2(r0, r1);
                *reinterpret_cast<float2*>(b0 + q) = make_float2((float)r0, (float)r1);
            } else {
                put2<TO, ACC>(out, q, z.re * scale, z.im * scale, b0);
            }
        }
        return;
    }
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) xl[s + 16 * k2] = v[4 * (k2 % 4) + k2 / 4];
    __syncthreads();
    C o[16][2];
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) {
        const int kk = s + 16 * k2;
        const cx<C> a = v[4 * (k2 % 4) + k2 / 4];
        if (kk == 0) {
            o[k2][0] = (a.re + a.im) * scale;   // x = 0
            o[k2][1] = (a.re - a.im) * scale;   // x = N - 1
        } else {
            const cx<C> F = r16::fwd_split(tw, kk, a, xl[M - kk]);
            o[k2][0] = F.re * scale;            // x = 2 kk - 1
            o[k2][1] = F.im * scale;            // x = 2 kk
        }
    }
    __syncthreads();
    // stage the line in LDS (reals) for aligned, coalesced stores
    C* fl = reinterpret_cast<C*>(xl);
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) {
        const int kk = s + 16 * k2;
        if (kk == 0) { fl[0] = o[k2][0]; fl[N - 1] = o[k2][1]; }
        else { fl[2 * kk - 1] = o[k2][0]; fl[2 * kk] = o[k2][1]; }
    }
    __syncthreads();
    if constexpr (std::is_same<C, float>::value) {
        // 16 lines x 128 float4
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int e = threadIdx.x + 256 * it, ln = e >> 7, c4 = e & 127;
            const float* src = reinterpret_cast<const float*>(xb + ln * LS) + 4 * c4;
            const long q = gidx(g, 4 * c4, (tile % nbl) * 16 + ln, k);
            *reinterpret_cast<float4*>(out + q) = make_float4(src[0], src[1], src[2], src[3]);
        }
    } else {
        // 16 lines x 256 double2
#pragma unroll
        for (int it = 0; it < 16; it++) {
            const int e = threadIdx.x + 256 * it, ln = e >> 8, c2 = e & 255;
            const double* src = reinterpret_cast<const double*>(xb + ln * LS) + 2 * c2;
            const long q = gidx(g, 2 * c2, (tile % nbl) * 16 + ln, k);
            *reinterpret_cast<double2*>(out + q) = make_double2(src[0], src[1]);
        }
    }
}

template <int DIR, int INV, class C, class TO, int ACC, int RES, class B>
static bool try_rfft512(Ctx& c, const C* tw, const C* in, TO* out, B* b0) {
    constexpr bool F32 = std::is_same<C, float>::value;
    constexpr bool F64 = std::is_same<C, double>::value && std::is_same<TO, double>::value && !RES && !ACC;
    if constexpr (!F32 && !F64) return false;
    else {
        const int N = DIR == 0 ? c.g.nx : c.g.ny;
        if (N != 512) return false;
        if (DIR == 1) {
            // along y the double exchange would need 256 KB of LDS: FP64 keeps k_rfft_n
            if constexpr (RES || ACC || !F32) return false;
            else {
                if (c.g.nx % 64) return false;
                hipLaunchKernelGGL((k_rfft512_y<INV, TO, ACC>), dim3((c.g.nx / 64) * c.g.nz), dim3(1024), 0, c.stream,
                                   c.g, tw, in, out);
                return true;
            }
        }
        if (c.g.ny % 16) return false;
        hipLaunchKernelGGL((k_rfft512_x<INV, C, TO, ACC, RES, B>), dim3((c.g.ny / 16) * c.g.nz), dim3(256), 0,
                           c.stream, c.g, tw, in, out, b0);
        return true;
    }
}

// the register kernels cover lines of 256 with 64 | nx (y) or 32 | ny (x): float, and
// double for the FP64 transforms (no residual, no accumulation)
template <int DIR, int INV, class C, class TO, int ACC, int RES, class B>
static bool try_rfft256(Ctx& c, const C* tw, const C* in, TO* out, B* b0) {
    constexpr bool F32 = std::is_same<C, float>::value;
    constexpr bool F64 = std::is_same<C, double>::value && std::is_same<TO, double>::value && !RES && !ACC;
    if constexpr (!F32 && !F64) return false;
    else {
        const int N = DIR == 0 ? c.g.nx : c.g.ny;
        if (N != 256) return false;
        if (DIR == 1) {
            if constexpr (RES || ACC) return false;
            else {
                if (c.g.nx % 64) return false;
                hipLaunchKernelGGL((k_rfft256_y<INV, C, TO, ACC>), dim3((c.g.nx / 64) * c.g.nz), dim3(512), 0,
                                   c.stream, c.g, tw, in, out);
                return true;
            }
        }
        if (c.g.ny % 32) return false;
        hipLaunchKernelGGL((k_rfft256_x<INV, C, TO, ACC, RES, B>), dim3((c.g.ny / 32) * c.g.nz), dim3(256), 0,
                           c.stream, c.g, tw, in, out, b0);
        return true;
    }
}

template <int DIR, int INV, int N, int LPB, class C, class TO, int ACC, int RES, class B>
static bool try_rfft_n(Ctx& c, const C* tw, const C* in, TO* out, int n, int nlines, B* b0) {
    if (n != N || nlines % LPB != 0) return false;
    hipLaunchKernelGGL((k_rfft_n<DIR, INV, N, LPB, C, TO, ACC, RES, B>), dim3((nlines / LPB) * c.g.nz),
                       dim3(RfftShape<N, LPB>::NTHR), 0, c.stream, c.g, tw, in, out, b0);
    return true;
}

// lines per block: along x 4; along y enough lines for 64-byte row segments
template <int DIR, int INV, class C, class TO, int ACC, int RES, class B>
static bool rfft_fixed(Ctx& c, const C* tw, const C* in, TO* out, int n, int nlines, B* b0) {
    constexpr int L = DIR == 0 ? 4 : (sizeof(C) == 4 ? 16 : 8);
    constexpr int L512 = L > 8 ? 8 : 4, L1024 = L > 8 ? 4 : 2;
    return try_rfft_n<DIR, INV, 32, L, C, TO, ACC, RES, B>(c, tw, in, out, n, nlines, b0) ||
           try_rfft_n<DIR, INV, 64, L, C, TO, ACC, RES, B>(c, tw, in, out, n, nlines, b0) ||
           try_rfft_n<DIR, INV, 128, L, C, TO, ACC, RES, B>(c, tw, in, out, n, nlines, b0) ||
           try_rfft_n<DIR, INV, 256, L, C, TO, ACC, RES, B>(c, tw, in, out, n, nlines, b0) ||
           try_rfft_n<DIR, INV, 512, L512, C, TO, ACC, RES, B>(c, tw, in, out, n, nlines, b0) ||
           try_rfft_n<DIR, INV, 1024, L1024, C, TO, ACC, RES, B>(c, tw, in, out, n, nlines, b0);
}

// Transform of every line along DIR.  With RES the forward x transform takes
// rhs - L b0 as its input (in = rhs); false means no fixed-size kernel covers the
// shape and the caller forms the residual itself.
template <int DIR, int INV, class C, class TO, int ACC, int RES = 0, class B = float>
static int rfft_launch(Ctx& c, const C* tw, const C* in, TO* out, B* b0 = nullptr) {
    const int N = DIR == 0 ? c.g.nx : c.g.ny;
    const int nlines = DIR == 0 ? c.g.ny : c.g.nx;
    if (try_rfft256<DIR, INV, C, TO, ACC, RES, B>(c, tw, in, out, b0) ||
        try_rfft512<DIR, INV, C, TO, ACC, RES, B>(c, tw, in, out, b0) ||
        rfft_fixed<DIR, INV, C, TO, ACC, RES, B>(c, tw, in, out, N, nlines, b0)) {
        HIP_OK(hipGetLastError());
        return ICLS_OK;
    }
    if (RES) return -1;
    // any other power of two: the runtime-sized kernel
    const int M = N / 2;
    int LPB = DIR == 0 ? 4 : 16;
    if (LPB > nlines) LPB = nlines;
    while (nlines % LPB) LPB >>= 1;
    while (LPB > 1 && (size_t)(2L * LPB * M + M) * sizeof(cx<C>) > 64 * 1024) LPB >>= 1;
    int TPL = M / 2 < 64 ? M / 2 : 64;
    if (TPL < 1) TPL = 1;
    while (LPB * TPL > 1024) TPL >>= 1;
    int nthr = LPB * TPL;
    if (nthr < 64) nthr = 64;
    size_t shm = (size_t)(2L * LPB * M + M) * sizeof(cx<C>);
    dim3 grd((nlines / LPB) * c.g.nz);
    hipLaunchKernelGGL((k_rfft<DIR, INV, C, TO, ACC, B>), grd, dim3(nthr), shm, c.stream, c.g, N, LPB, TPL, tw, in, out, b0);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// Thomas elimination per (i,j) column (solver.f90:119-160) with the Periodic-branch
// matrix of initsolver.f90:36-49: a = c = dzi^2, b = -(a+c), b(1) += a, b(kmax) += c,
// a(1) = c(kmax) = 0, diagonal shifted by xyrt(i,j).  A zero last pivot (the (0,0)
// mode) gives p(kmax) = 0 as in the reference.  The arithmetic, the multipliers d
// and the forward-sweep values y are double; the right-hand side and the solution are
// stored as TS.  When TS is double, y is p itself (the reference's in-place sweep).
template <class TS>
__global__ void __launch_bounds__(64) k_thomas(Grid g, const double* __restrict__ xyrt, TS* p, double* __restrict__ d, double* y) {
    int i0 = blockIdx.x * blockDim.x + threadIdx.x, j0 = blockIdx.y;
    if (i0 >= g.nx) return;
    const int kmax = g.nz;
    const double a = g.dzi * g.dzi, cc = g.dzi * g.dzi;
    const double bmid = -(a + cc);
    const double b1 = bmid + a, bN = bmid + cc;
    const double lam = xyrt[i0 + (long)g.nx * j0];
    const long sz = g.sz;
    const long q1 = gidx(g, i0, j0, 1);
    // The recurrences are serial in k; the loads they consume are not, so each sweep
    // fetches a chunk of CH values ahead of the arithmetic that uses them.  Full
    // chunks run without bounds checks so every register index is static; a scalar
    // tail takes the remainder.
    constexpr int CH = 32;
    double dprev = 0., pprev = 0.;
    // one forward step at layer k with right-hand side rk; returns y(k), sets d(k)
    auto fwd = [&](int k, double rk) {
        double z;
        if (k == 1) {
            z = 1. / (b1 + lam);
            dprev = cc * z;
            pprev = rk * z;
        } else if (k < kmax) {
            double bb = bmid + lam;
            z = 1. / (bb - a * dprev);
            dprev = cc * z;
            pprev = (rk - a * pprev) * z;
        } else {
            double bb = bN + lam;
            z = bb - a * dprev;
            if (z != 0.) pprev = (rk - a * pprev) / z;
            else pprev = 0.;
        }
    };
    auto put_fwd = [&](int k, double yk, double dk) {
        const long q = q1 + (long)(k - 1) * sz;
        if (k < kmax) { y[q] = yk; d[q] = dk; }
        else p[q] = (TS)yk;   // p(kmax) is final
    };
    int kb = 1;
    for (; kb + CH - 1 <= kmax; kb += CH) {
        double buf[CH], dbuf[CH];
#pragma unroll
        for (int t = 0; t < CH; t++) buf[t] = (double)p[q1 + (long)(kb + t - 1) * sz];
#pragma unroll
        for (int t = 0; t < CH; t++) {
            fwd(kb + t, buf[t]);
            buf[t] = pprev;
            dbuf[t] = dprev;
        }
#pragma unroll
        for (int t = 0; t < CH; t++) put_fwd(kb + t, buf[t], dbuf[t]);
    }
    for (; kb <= kmax; kb++) {
        fwd(kb, (double)p[q1 + (long)(kb - 1) * sz]);
        put_fwd(kb, pprev, dprev);
    }
    // back substitution from k = kmax-1 down to 1; pprev holds p(kmax)
    int kt = kmax - 1;
    for (; kt - CH + 1 >= 1; kt -= CH) {
        double buf[CH], dbuf[CH];
#pragma unroll
        for (int t = 0; t < CH; t++) {
            const long q = q1 + (long)(kt - t - 1) * sz;
            buf[t] = y[q];
            dbuf[t] = d[q];
        }
#pragma unroll
        for (int t = 0; t < CH; t++) {
            pprev = buf[t] - dbuf[t] * pprev;
            buf[t] = pprev;
        }
#pragma unroll
        for (int t = 0; t < CH; t++) p[q1 + (long)(kt - t - 1) * sz] = (TS)buf[t];
    }
    for (; kt >= 1; kt--) {
        const long q = q1 + (long)(kt - 1) * sz;
        pprev = y[q] - d[q] * pprev;
        p[q] = (TS)pprev;
    }
}

// The same elimination with every column cut into P = nz / S segments of S layers.
// Wave s of a block holds segment s of 64 consecutive columns in registers, so the
// multipliers and the forward-sweep values never leave the chip.
//
// The multipliers d(k) do not depend on the right-hand side: each segment restarts
// their recurrence from d at the layer below it (dseed, computed once per grid by
// k_thomas_seed with the same operations), so every d(k) and 1/pivot is bit for bit
// the serial one.  The forward sweep y(k) = (r(k) - a y(k-1)) / pivot(k) is affine in
// the value entering a segment: a first pass over the segment with y = 0 entering
// gives its exit value and the coefficient of the entering value; the carries are
// folded across segments through LDS, and a second pass repeats the serial
// recurrence from the true entering value.  The back substitution
// p(k) = y(k) - d(k) p(k+1) is treated the same way from the top of the segment.
// Only the entering values differ from the serial sweep, by rounding.
template <int S>
struct ThomasRow {
    double a, cc, bmid, b1, bN, lam;
    int n;
    // 1 / pivot at layer k (k < n), or the pivot itself at k = n; advances d
    __device__ __forceinline__ double piv(int k, double& dprev) const {
        double z;
        if (k == 1) {
            z = 1. / (b1 + lam);
            dprev = cc * z;
        } else if (k < n) {
            double bb = bmid + lam;
            z = 1. / (bb - a * dprev);
            dprev = cc * z;
        } else {
            double bb = bN + lam;
            z = bb - a * dprev;
        }
        return z;
    }
    __device__ __forceinline__ double fwd(int k, double rk, double z, double yprev) const {
        if (k == 1) return rk * z;
        if (k < n) return (rk - a * yprev) * z;
        return z != 0. ? (rk - a * yprev) / z : 0.;
    }
};

__global__ void k_thomas_seed(Grid g, const double* __restrict__ xyrt, double* __restrict__ dseed, int S) {
    const long ncol = (long)g.nx * g.ny;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    ThomasRow<1> R;
    R.a = g.dzi * g.dzi;
    R.cc = g.dzi * g.dzi;
    R.bmid = -(R.a + R.cc);
    R.b1 = R.bmid + R.a;
    R.bN = R.bmid + R.cc;
    R.lam = xyrt[c];
    R.n = g.nz;
    double dprev = 0.;
    for (int k = 1; k < g.nz; k++) {
        R.piv(k, dprev);
        if (k % S == 0) dseed[(long)(k / S - 1) * ncol + c] = dprev;
    }
}

// CPW columns per wave: with 64 a wave is one segment of 64 columns; with 32 a wave
// holds two segments of 32 columns (fewer registers per lane at S = 8, so that two
// blocks fit a CU).  A block is P segments x CPW columns.
#ifndef ICLS_THOMAS_CPW
#define ICLS_THOMAS_CPW 32
#endif
constexpr int TCPW = ICLS_THOMAS_CPW;
constexpr int TMAXSEG = 16 * (64 / TCPW);   // segments per block at 1024 threads

template <class TS, int S>
__global__ void __launch_bounds__(1024, (TCPW == 32 && S <= 8 ? 8 : 4))
k_thomas_seg(Grid g, const double* __restrict__ xyrt, const double* __restrict__ dseed, TS* p, int ngrp) {
    __shared__ double cv[TMAXSEG][TCPW], cf[TMAXSEG][TCPW];   // segment carries: value, coefficient
    const int lane = threadIdx.x % TCPW, s = threadIdx.x / TCPW;
    const int P = blockDim.x / TCPW;
    const long ncol = (long)g.nx * g.ny;
    ThomasRow<S> R;
    R.a = g.dzi * g.dzi;
    R.cc = g.dzi * g.dzi;
    R.bmid = -(R.a + R.cc);
    R.b1 = R.bmid + R.a;
    R.bN = R.bmid + R.cc;
    R.n = g.nz;
    const int k0 = s * S + 1;
    const long qk = gidx(g, 0, 0, k0);
    // A block takes the column groups blockIdx.x, + gridDim.x, ...  With PF the
    // right-hand side, the mode's shift and the multiplier seed of the next group are
    // loaded before the current group is eliminated, so the loads are in flight during
    // the arithmetic (where the registers allow it: float storage, 16 layers).
    constexpr bool PF = sizeof(TS) == 4 && S == 16;
    TS rn[S];
    double lam_n = 0., seed_n = 0.;
    auto fetch = [&](int grp) {
        const long c = (long)grp * TCPW + lane;
        const bool act = c < ncol;
#pragma unroll
        for (int t = 0; t < S; t++) rn[t] = act ? p[qk + c + (long)t * g.sz] : TS(0.);
        lam_n = act ? xyrt[c] : 0.;
        seed_n = (s > 0 && act) ? dseed[(long)(s - 1) * ncol + c] : 0.;
    };
    int grp = blockIdx.x;
    if (PF && grp < ngrp) fetch(grp);
    for (; grp < ngrp; grp += gridDim.x) {
        if (!PF) fetch(grp);
        const long c = (long)grp * TCPW + lane;
        const bool act = c < ncol;
        const long q0 = qk + c;
        double y[S], z[S];
#pragma unroll
        for (int t = 0; t < S; t++) y[t] = (double)rn[t];
        R.lam = lam_n;
        double dprev = seed_n;
        if (PF && grp + (int)gridDim.x < ngrp) fetch(grp + gridDim.x);
#pragma unroll
        for (int t = 0; t < S; t++) z[t] = R.piv(k0 + t, dprev);
        // forward: exit value with 0 entering, and the coefficient of the entering value
        {
            double v = 0., f = 1.;
#pragma unroll
            for (int t = 0; t < S; t++) {
                const int k = k0 + t;
                v = R.fwd(k, y[t], z[t], v);
                if (k == 1) f = 0.;
                else if (k < R.n) f = -R.a * z[t] * f;
                else f = z[t] != 0. ? -R.a * f / z[t] : 0.;
            }
            cv[s][lane] = v;
            cf[s][lane] = f;
        }
        __syncthreads();
        double yin = 0.;
        for (int t = 0; t < s; t++) yin = cv[t][lane] + cf[t][lane] * yin;
#pragma unroll
        for (int t = 0; t < S; t++) {
            yin = R.fwd(k0 + t, y[t], z[t], yin);
            y[t] = yin;
        }
        __syncthreads();   // the carries are read before they are replaced
        // back substitution: value at the segment's first layer with 0 above it, and
        // the coefficient of the value above
        {
            double v = 0., f = 1.;
#pragma unroll
            for (int t = S - 1; t >= 0; t--) {
                const int k = k0 + t;
                if (k == R.n) { v = y[t]; f = 0.; }
                else {
                    const double d = R.cc * z[t];
                    v = y[t] - d * v;
                    f = -d * f;
                }
            }
            cv[s][lane] = v;
            cf[s][lane] = f;
        }
        __syncthreads();
        double pa = 0.;
        for (int t = P - 1; t > s; t--) pa = cv[t][lane] + cf[t][lane] * pa;
#pragma unroll
        for (int t = S - 1; t >= 0; t--) {
            const int k = k0 + t;
            if (k == R.n) pa = y[t];
            else pa = y[t] - (R.cc * z[t]) * pa;
            y[t] = pa;
        }
        if (act)
#pragma unroll
            for (int t = 0; t < S; t++) p[q0 + (long)t * g.sz] = (TS)y[t];
        __syncthreads();   // the carries are read before the next group writes them
    }
}

// Segment length of the partitioned solve for nz layers, 0 where the serial kernel
// runs (a block holds at most TMAXSEG segments).
int thomas_seg_len(int nz) {
    if (nz % 16 == 0 && nz / 16 <= TMAXSEG && nz / 16 >= 64 / TCPW) return 16;
    if (nz % 8 == 0 && nz / 8 <= TMAXSEG) return 8;
    return 0;
}

int launch_thomas_seed(Ctx& c) {
    if (!c.thomas_seg) return ICLS_OK;
    const long ncol = (long)c.g.nx * c.g.ny;
    hipLaunchKernelGGL(k_thomas_seed, dim3((unsigned)((ncol + 63) / 64)), dim3(64), 0, c.stream, c.g,
                       (const double*)c.xyrt, c.thomas_seed, c.thomas_seg);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// blocks of a looping kernel: as many as are resident at once, no more than there is
// work for (two rounds of resident blocks measured the same as one)
static unsigned looping_grid(Ctx& c, const void* fn, int threads, long nwork) {
    auto it = c.resident.find(fn);
    int occ;
    if (it != c.resident.end()) occ = it->second;
    else {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, threads, 0) != hipSuccess || occ < 1) occ = 1;
        c.resident[fn] = occ;
    }
    const long nb = (long)occ * (c.ncu > 0 ? c.ncu : 1);
    return (unsigned)(nb < nwork ? nb : nwork);
}

template <class TS>
static int thomas_launch(Ctx& c, TS* f, double* y) {
    ProfScope ps(c, "k_thomas");
    if (c.thomas_seg) {
        const long ncol = (long)c.g.nx * c.g.ny;
        const int P = c.g.nz / c.thomas_seg;
        const int ngrp = (int)((ncol + TCPW - 1) / TCPW);
        dim3 blk(TCPW * P);
        if (c.thomas_seg == 16) {
            const void* fn = (const void*)k_thomas_seg<TS, 16>;
            hipLaunchKernelGGL((k_thomas_seg<TS, 16>), dim3(looping_grid(c, fn, blk.x, ngrp)), blk, 0, c.stream, c.g,
                               (const double*)c.xyrt, (const double*)c.thomas_seed, f, ngrp);
        } else {
            const void* fn = (const void*)k_thomas_seg<TS, 8>;
            hipLaunchKernelGGL((k_thomas_seg<TS, 8>), dim3(looping_grid(c, fn, blk.x, ngrp)), blk, 0, c.stream, c.g,
                               (const double*)c.xyrt, (const double*)c.thomas_seed, f, ngrp);
        }
        HIP_OK(hipGetLastError());
        return ICLS_OK;
    }
    int bx = c.g.nx < 64 ? c.g.nx : 64;
    hipLaunchKernelGGL(k_thomas<TS>, dim3((c.g.nx + bx - 1) / bx, c.g.ny, 1), dim3(bx), 0, c.stream, c.g,
                       (const double*)c.xyrt, f, c.thomas_d, y);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// the residual as an array, for shapes the fused forward transform does not cover
template <class T, class B>
__global__ void k_residual(Grid g, const T* __restrict__ rhs, const B* __restrict__ p, T* __restrict__ r) {
    int i0 = blockIdx.x * 64 + threadIdx.x, j0 = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    r[gidx(g, i0, j0, k)] = (T)residual_at(g, rhs, p, i0, j0, k);
}

// ---------------------------------------------------------------- the solve
static int solve_fp64(Ctx& c) {
    const double* twx = c.tw_x;
    const double* twy = c.tw_y;
    double* p = c.p;
    {
        ProfScope ps(c, "k_rfft_x_fwd");
        int rc = rfft_launch<0, 0, double, double, 0>(c, twx, p, p);
        if (rc) return rc;
    }
    {
        ProfScope ps(c, "k_rfft_y_fwd");
        int rc = rfft_launch<1, 0, double, double, 0>(c, twy, p, p);
        if (rc) return rc;
    }
    int rc = thomas_launch<double>(c, p, p);
    if (rc) return rc;
    {
        ProfScope ps(c, "k_rfft_y_inv");
        rc = rfft_launch<1, 1, double, double, 0>(c, twy, p, p);
        if (rc) return rc;
    }
    ProfScope ps(c, "k_rfft_x_inv");
    return rfft_launch<0, 1, double, double, 0>(c, twx, p, p);
}

// One float-transform solve.  ACC = 0: of L p0 = rhs, the solution written as float to
// p0 (the transforms produce float values, so nothing is lost against a double
// store).  ACC = 1: of L e = rhs - L p0 for the float starting solution p0, the
// residual formed in double inside the first transform where a fixed-size kernel
// covers the line length and by k_residual otherwise; the last transform writes
// p = p0 + e as double and p rounded to float into p0.
template <int ACC>
static int solve_fp32_pass(Ctx& c) {
    const float* twx = as<float>(c.tw_x_f);
    const float* twy = as<float>(c.tw_y_f);
    const float* rhs = as<float>(c.frhs);
    float* w = as<float>(c.pf);
    float* p0 = as<float>(c.pf0);
    int rc;
    if (!ACC) {
        ProfScope ps(c, "k_rfft_x_fwd");
        rc = rfft_launch<0, 0, float, float, 0>(c, twx, rhs, w);
        if (rc) return rc;
    } else {
        ProfScope ps(c, "k_rfft_x_fwd_res");
        rc = rfft_launch<0, 0, float, float, 0, 1>(c, twx, rhs, w, p0);
        if (rc < 0) {
            dim3 blk(64, 4, 1);
            hipLaunchKernelGGL((k_residual<float, float>), grid3(c.g, c.g.nz, blk), blk, 0, c.stream, c.g, rhs,
                               (const float*)p0, w);
            HIP_OK(hipGetLastError());
            rc = rfft_launch<0, 0, float, float, 0>(c, twx, w, w);
        }
        if (rc) return rc;
    }
    {
        ProfScope ps(c, "k_rfft_y_fwd");
        rc = rfft_launch<1, 0, float, float, 0>(c, twy, w, w);
        if (rc) return rc;
    }
    rc = thomas_launch<float>(c, w, c.thomas_y);
    if (rc) return rc;
    {
        ProfScope ps(c, "k_rfft_y_inv");
        rc = rfft_launch<1, 1, float, float, 0>(c, twy, w, w);
        if (rc) return rc;
    }
    ProfScope ps(c, "k_rfft_x_inv");
    if (!ACC) return rfft_launch<0, 1, float, float, 0>(c, twx, w, p0);
    return rfft_launch<0, 1, float, double, 1>(c, twx, w, c.p, p0);
}

// One step of iterative refinement from the float starting solution p0 (pf0).  Inside
// the time step p0 is, where there is one, the previous step's solution rounded to
// float, which the last refinement left in pf0: one pass then solves for the change of
// the pressure over the step (and the float rounding of the start), whose float error
// is that of the change.  Otherwise (the first step, after a restart or an upload of
// the pressures, after a solve through the API, and for every API solve) a float solve
// provides p0 first.
static int solve_mixed(Ctx& c, bool from_previous) {
    if (!from_previous) {
        int rc = solve_fp32_pass<0>(c);
        if (rc) return rc;
    }
    return solve_fp32_pass<1>(c);
}

int launch_solve(Ctx& c, bool warm) {
    if (!c.fp32) return solve_fp64(c);
    // under a lazy rotation the pnew buffer holds the raw solution of the last solve,
    // and pf0 that solution rounded to float when that solve was the time step's
    const bool from_previous = warm && c.p_lazy && c.pf0_step;
    const int rc = solve_mixed(c, from_previous);
    c.pf0_step = warm && rc == ICLS_OK;
    return rc;
}

// ---------------------------------------------------------------- solver2d_cos
// solver.f90:202-317 with initsolver.f90:84-160: real FFT along x, the cosine transform
// VCOSQB along z, Gauss elimination along y with the boundary rule of BC_in_y,
// VCOSQF along z, inverse FFT along x.  On one GPU the transposes are identities.
//
// Both z transforms are dense products over the nz layers of every (i, j) column,
// out(col, kk) = sum_k in(col, k) M(k, kk), with M the matrix of the transform (fft.f:
// 116-120, 266-270; VFFTPACK's fast algorithm computes the same sums in another order).
// A block computes 64 columns x 64 outputs, staging 16 layers of the input and of M in
// LDS per pass; each thread holds 4 x 4 sums (columns tx + 16 r, outputs ty + 16 s).
constexpr int ZT_COLS = 64, ZT_OUT = 64, ZT_K = 16;

__global__ void __launch_bounds__(256) k_ztransform(Grid g, const double* __restrict__ M,
                                                    const double* __restrict__ in, double* __restrict__ out) {
    __shared__ double As[ZT_K][ZT_COLS];
    __shared__ double Bs[ZT_K][ZT_OUT];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const long ncol = (long)g.nx * g.ny;
    const long col0 = (long)blockIdx.x * ZT_COLS;
    const int kk0 = blockIdx.y * ZT_OUT;
    const int nz = g.nz;
    const double* in0 = in + g.sz * 3;   // layer k = 1
    double acc[4][4] = {};
    for (int kb = 0; kb < nz; kb += ZT_K) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = threadIdx.x + 256 * q, kq = e >> 6, cc = e & 63;
            const long col = col0 + cc;
            As[kq][cc] = (kb + kq < nz && col < ncol) ? in0[col + g.sz * (kb + kq)] : 0.;
            Bs[kq][cc] = (kb + kq < nz && kk0 + cc < nz) ? M[(long)(kb + kq) * nz + kk0 + cc] : 0.;
        }
        __syncthreads();
#pragma unroll
        for (int kq = 0; kq < ZT_K; kq++) {
            double av[4], bv[4];
#pragma unroll
            for (int r = 0; r < 4; r++) av[r] = As[kq][tx + 16 * r];
#pragma unroll
            for (int s = 0; s < 4; s++) bv[s] = Bs[kq][ty + 16 * s];
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int s = 0; s < 4; s++) acc[r][s] = fma(av[r], bv[s], acc[r][s]);
        }
        __syncthreads();
    }
    double* out0 = out + g.sz * 3;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int kk = kk0 + ty + 16 * s;
        if (kk >= nz) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const long col = col0 + tx + 16 * r;
            if (col < ncol) out0[col + g.sz * kk] = acc[r][s];
        }
    }
}

// The same products with half the work for even nz = 2H, from the symmetry of the two
// matrices (0-based indices): M_b(nz-1-k, i) = (-1)^i M_b(k, i) and M_f(k, nz-1-i) =
// (-1)^k M_f(k, i).
//   VCOSQB: out(2j)   = sum_{k<H} (x(k) + x(nz-1-k)) M_b(k, 2j)
//           out(2j+1) = sum_{k<H} (x(k) - x(nz-1-k)) M_b(k, 2j+1)
//   VCOSQF: E(i) = sum_{m<H} x(2m) M_f(2m, i),  O(i) = sum_{m<H} x(2m+1) M_f(2m+1, i),
//           out(i) = E(i) + O(i),  out(nz-1-i) = E(i) - O(i)   (i < H)
// P and Q are the two H x H matrices (row = m, contiguous j or i): M_b(m, 2j) and
// M_b(m, 2j+1) forward, M_f(2m, i) and M_f(2m+1, i) inverse.  A block computes 64
// columns x 64 indices j (or i) of both halves.
template <int INV>
__global__ void __launch_bounds__(256) k_ztransform_sym(Grid g, const double* __restrict__ P,
                                                        const double* __restrict__ Q, const double* __restrict__ in,
                                                        double* __restrict__ out) {
    __shared__ double A0[ZT_K][ZT_COLS], A1[ZT_K][ZT_COLS];
    __shared__ double B0[ZT_K][ZT_OUT], B1[ZT_K][ZT_OUT];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const long ncol = (long)g.nx * g.ny;
    const long col0 = (long)blockIdx.x * ZT_COLS;
    const int j0 = blockIdx.y * ZT_OUT;
    const int nz = g.nz, H = nz / 2;
    const double* in0 = in + g.sz * 3;   // layer k = 1
    double e[4][4] = {}, o[4][4] = {};
    for (int mb = 0; mb < H; mb += ZT_K) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int el = threadIdx.x + 256 * q, mq = el >> 6, cc = el & 63;
            const int m = mb + mq;
            const long col = col0 + cc;
            double a0 = 0., a1 = 0.;
            if (m < H && col < ncol) {
                if (INV) {
                    a0 = in0[col + g.sz * (2 * m)];
                    a1 = in0[col + g.sz * (2 * m + 1)];
                } else {
                    const double xl = in0[col + g.sz * m], xh = in0[col + g.sz * (nz - 1 - m)];
                    a0 = xl + xh;
                    a1 = xl - xh;
                }
            }
            A0[mq][cc] = a0;
            A1[mq][cc] = a1;
            const bool okb = m < H && j0 + cc < H;
            B0[mq][cc] = okb ? P[(long)m * H + j0 + cc] : 0.;
            B1[mq][cc] = okb ? Q[(long)m * H + j0 + cc] : 0.;
        }
        __syncthreads();
#pragma unroll
        for (int mq = 0; mq < ZT_K; mq++) {
            double a0[4], a1[4], b0[4], b1[4];
#pragma unroll
            for (int r = 0; r < 4; r++) { a0[r] = A0[mq][tx + 16 * r]; a1[r] = A1[mq][tx + 16 * r]; }
#pragma unroll
            for (int t = 0; t < 4; t++) { b0[t] = B0[mq][ty + 16 * t]; b1[t] = B1[mq][ty + 16 * t]; }
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    e[r][t] = fma(a0[r], b0[t], e[r][t]);
                    o[r][t] = fma(a1[r], b1[t], o[r][t]);
                }
        }
        __syncthreads();
    }
    double* out0 = out + g.sz * 3;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int j = j0 + ty + 16 * t;
        if (j >= H) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const long col = col0 + tx + 16 * r;
            if (col >= ncol) continue;
            if (INV) {
                out0[col + g.sz * j] = e[r][t] + o[r][t];
                out0[col + g.sz * (nz - 1 - j)] = e[r][t] - o[r][t];
            } else {
                out0[col + g.sz * (2 * j)] = e[r][t];
                out0[col + g.sz * (2 * j + 1)] = o[r][t];
            }
        }
    }
}

// Gauss elimination along y per (i, k) line (solver.f90:253-290): a = c = dyi^2,
// b = -(a + c), the end rows changed by BC_in_y (initsolver.f90:88-124; 1 Walls:
// b(1) += a(1), b(jtot) += c(jtot); 2 InOut: b(1) += a(1), b(jtot) -= c(jtot);
// 3 OutOut: b(1) -= a(1), b(jtot) -= c(jtot)), then a(1) = c(jtot) = 0, the diagonal
// shifted by xrt(i) + zrt(k); a zero last pivot gives p(jtot) = 0.  The sweep runs in
// place on q (the transformed right-hand side); the multipliers go to dst.
__global__ void __launch_bounds__(64) k_thomas_ycos(Grid g, int bcy, const double* __restrict__ xrt,
                                                    const double* __restrict__ zrt, double* q,
                                                    double* __restrict__ dst) {
    const int i0 = blockIdx.x * 64 + threadIdx.x, k = blockIdx.y + 1;
    if (i0 >= g.nx) return;
    const int jmax = g.ny;
    const double a = g.dyi * g.dyi, cc = g.dyi * g.dyi;
    const double bmid = -(a + cc);
    const double b1 = bcy == 3 ? bmid - a : bmid + a;
    const double bN = bcy == 1 ? bmid + cc : bmid - cc;
    const double xz = xrt[i0 + 1] + zrt[k];
    const long sy = g.sy;
    const long q1 = gidx(g, i0, 0, k);
    constexpr int CH = 32;
    double dprev = 0., pprev = 0.;
    auto fwd = [&](int j, double rj) {
        double z;
        if (j == 1) {
            z = 1. / (b1 + xz);
            dprev = cc * z;
            pprev = rj * z;
        } else if (j < jmax) {
            double bb = bmid + xz;
            z = 1. / (bb - a * dprev);
            dprev = cc * z;
            pprev = (rj - a * pprev) * z;
        } else {
            double bb = bN + xz;
            z = bb - a * dprev;
            if (z != 0.) pprev = (rj - a * pprev) / z;
            else pprev = 0.;
        }
    };
    int jb = 1;
    for (; jb + CH - 1 <= jmax; jb += CH) {
        double buf[CH], dbuf[CH];
#pragma unroll
        for (int t = 0; t < CH; t++) buf[t] = q[q1 + (long)(jb + t - 1) * sy];
#pragma unroll
        for (int t = 0; t < CH; t++) {
            fwd(jb + t, buf[t]);
            buf[t] = pprev;
            dbuf[t] = dprev;
        }
#pragma unroll
        for (int t = 0; t < CH; t++) {
            const long o = q1 + (long)(jb + t - 1) * sy;
            q[o] = buf[t];
            dst[o] = dbuf[t];
        }
    }
    for (; jb <= jmax; jb++) {
        const long o = q1 + (long)(jb - 1) * sy;
        fwd(jb, q[o]);
        q[o] = pprev;
        dst[o] = dprev;
    }
    // back substitution from j = jmax-1 down to 1; pprev holds p(jmax)
    int jt = jmax - 1;
    for (; jt - CH + 1 >= 1; jt -= CH) {
        double buf[CH], dbuf[CH];
#pragma unroll
        for (int t = 0; t < CH; t++) {
            const long o = q1 + (long)(jt - t - 1) * sy;
            buf[t] = q[o];
            dbuf[t] = dst[o];
        }
#pragma unroll
        for (int t = 0; t < CH; t++) {
            pprev = buf[t] - dbuf[t] * pprev;
            buf[t] = pprev;
        }
#pragma unroll
        for (int t = 0; t < CH; t++) q[q1 + (long)(jt - t - 1) * sy] = buf[t];
    }
    for (; jt >= 1; jt--) {
        const long o = q1 + (long)(jt - 1) * sy;
        pprev = q[o] - dst[o] * pprev;
        q[o] = pprev;
    }
}

// tables of solver2d_cos, built on the host once per context: the VCOSQB and VCOSQF
// matrices (row = input layer, contiguous outputs) and xrt(1..nx), zrt(1..nz)
// (initsolver.f90:132-145)
static int cos_tables(Ctx& c) {
    if (c.cos_tab) return ICLS_OK;
    const int nx = c.g.nx, nz = c.g.nz, H = nz / 2;
    const double pi = c.pi, s = sqrt(4. * nz);
    std::vector<double> t((size_t)2 * nz * nz + (nx + 1) + (nz + 1) + (size_t)4 * H * H, 0.);
    double* mb = t.data();
    double* mf = mb + (size_t)nz * nz;
    double* xrt = mf + (size_t)nz * nz;
    double* zrt = xrt + nx + 1;
    double* half = zrt + nz + 1;   // the halves of the even-nz products: P_b, Q_b, P_f, Q_f
    for (int k = 1; k <= nz; k++)
        for (int i = 1; i <= nz; i++) {
            mb[(size_t)(k - 1) * nz + (i - 1)] = 4. * cos((double)((2 * k - 1) * (i - 1)) * pi / (2. * nz)) / s;
            mf[(size_t)(k - 1) * nz + (i - 1)] =
                (k == 1 ? 1. : 2. * cos((double)((2 * i - 1) * (k - 1)) * pi / (2. * nz))) / s;
        }
    for (int i = 3; i <= nx; i += 2) {
        double sn = sin((double)(i - 1) * pi / (2. * nx));
        xrt[i - 1] = -4. * c.g.dxi * c.g.dxi * (sn * sn);
        xrt[i] = xrt[i - 1];
    }
    xrt[1] = 0.;
    xrt[nx] = -4. * c.g.dxi * c.g.dxi;
    for (int k = 1; k <= nz; k++) zrt[k] = (c.g.dzi * c.g.dzi) * (-2. + 2. * cos((double)(k - 1) * pi / nz));
    for (int m = 0; m < H; m++)
        for (int j = 0; j < H; j++) {
            half[(size_t)m * H + j] = mb[(size_t)m * nz + 2 * j];
            half[(size_t)H * H + (size_t)m * H + j] = mb[(size_t)m * nz + 2 * j + 1];
            half[(size_t)2 * H * H + (size_t)m * H + j] = mf[(size_t)(2 * m) * nz + j];
            half[(size_t)3 * H * H + (size_t)m * H + j] = mf[(size_t)(2 * m + 1) * nz + j];
        }
    HIP_OK(hipMalloc((void**)&c.cos_tab, sizeof(double) * t.size()));
    HIP_OK(hipMemcpy(c.cos_tab, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
    return ICLS_OK;
}

int launch_solve_cos(Ctx& c, int bcy) {
    if (bcy < 1 || bcy > 3) return ICLS_ERR_PARAM;
    if (c.g.ny < 2) return ICLS_ERR_PARAM;
    int rc = cos_tables(c);
    if (rc) return rc;
    const int nz = c.g.nz;
    const double* mb = c.cos_tab;
    const double* mf = mb + (size_t)nz * nz;
    const double* xrt = mf + (size_t)nz * nz;
    const double* zrt = xrt + c.g.nx + 1;
    const int H = nz / 2;
    const double* half = zrt + nz + 1;
    // the halved products from nz = 128 on: at 256^3 they took 196 / 227 us against
    // 347 / 352 us; at nz = 64 (H = 32, half of each 64-wide tile idle) 106 / 126 us
    // against 99 / 103 us
    const bool sym = nz % 2 == 0 && nz >= 128;
    double* p = c.p;
    double* w = c.scratch;
    {
        ProfScope ps(c, "k_rfft_x_fwd");
        rc = rfft_launch<0, 0, double, double, 0>(c, c.tw_x, p, p);
        if (rc) return rc;
    }
    const long ncol = (long)c.g.nx * c.g.ny;
    dim3 zg((unsigned)((ncol + ZT_COLS - 1) / ZT_COLS), (unsigned)((nz + ZT_OUT - 1) / ZT_OUT));
    dim3 zh((unsigned)((ncol + ZT_COLS - 1) / ZT_COLS), (unsigned)((H + ZT_OUT - 1) / ZT_OUT));
    {
        ProfScope ps(c, "k_ztransform_fwd");
        if (sym)
            hipLaunchKernelGGL(k_ztransform_sym<0>, zh, dim3(256), 0, c.stream, c.g, half, half + (size_t)H * H,
                               (const double*)p, w);
        else
            hipLaunchKernelGGL(k_ztransform, zg, dim3(256), 0, c.stream, c.g, mb, (const double*)p, w);
        HIP_OK(hipGetLastError());
    }
    {
        ProfScope ps(c, "k_thomas_ycos");
        hipLaunchKernelGGL(k_thomas_ycos, dim3((c.g.nx + 63) / 64, nz), dim3(64), 0, c.stream, c.g, bcy, xrt, zrt,
                           w, p);
        HIP_OK(hipGetLastError());
    }
    {
        ProfScope ps(c, "k_ztransform_inv");
        if (sym)
            hipLaunchKernelGGL(k_ztransform_sym<1>, zh, dim3(256), 0, c.stream, c.g, half + (size_t)2 * H * H,
                               half + (size_t)3 * H * H, (const double*)w, p);
        else
            hipLaunchKernelGGL(k_ztransform, zg, dim3(256), 0, c.stream, c.g, mf, (const double*)w, p);
        HIP_OK(hipGetLastError());
    }
    ProfScope ps(c, "k_rfft_x_inv");
    return rfft_launch<0, 1, double, double, 0>(c, c.tw_x, p, p);
}

}  // namespace icls
