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
#include "icls_internal.hpp"

namespace icls {

template <class C> struct cx { C re, im; };
template <class C> __device__ __forceinline__ cx<C> cmul(cx<C> a, cx<C> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class C> __device__ __forceinline__ cx<C> cadd(cx<C> a, cx<C> b) { return {a.re + b.re, a.im + b.im}; }
template <class C> __device__ __forceinline__ cx<C> csub(cx<C> a, cx<C> b) { return {a.re - b.re, a.im - b.im}; }
template <class C> __device__ __forceinline__ cx<C> cconj(cx<C> a) { return {a.re, -a.im}; }

template <class TO, class C, int ACC>
__device__ __forceinline__ void put(TO* out, long q, C v) {
    if (ACC) out[q] = out[q] + (TO)v;
    else out[q] = (TO)v;
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
template <int DIR, int INV, class C, class TO, int ACC>
__global__ void k_rfft(Grid g, int N, int LPB, int TPL, const C* __restrict__ twg, const C* in, TO* out) {
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
            put<TO, C, ACC>(out, gq(lnn, n), scale * v);
        }
    } else {
        for (int e = tid; e < LPB * N; e += nthr) {
            int lnn, n;
            split(e, lnn, n);
            cx<C> z = outb[(long)lnn * M + (n >> 1)];
            put<TO, C, ACC>(out, gq(lnn, n), scale * ((n & 1) ? z.im : z.re));
        }
    }
}

// The same transform with the line length N and the lines per block LPB fixed at
// compile time: every index split is a shift and every stage loop unrolls.
template <int N, int LPB>
struct RfftShape {
    static constexpr int M = N / 2;
    static constexpr int TPL = (M / 2) < 64 ? (M / 2) : 64;
    static constexpr int NTHR = LPB * TPL < 64 ? 64 : LPB * TPL;
};

template <int DIR, int INV, int N, int LPB, class C, class TO, int ACC>
__global__ void __launch_bounds__((RfftShape<N, LPB>::NTHR))
k_rfft_n(Grid g, const C* __restrict__ twg, const C* in, TO* out) {
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
            C v = in[gq(ln, n)];
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
            put<TO, C, ACC>(out, gq(lnn, n), scale * v);
        }
    } else {
#pragma unroll 4
        for (int e = tid; e < LPB * N; e += NTHR) {
            int lnn, n;
            split(e, lnn, n);
            cx<C> z = outb[lnn * M + (n >> 1)];
            put<TO, C, ACC>(out, gq(lnn, n), scale * ((n & 1) ? z.im : z.re));
        }
    }
}

template <int DIR, int INV, int N, int LPB, class C, class TO, int ACC>
static bool try_rfft_n(Ctx& c, const C* tw, const C* in, TO* out, int n, int nlines) {
    if (n != N || nlines % LPB != 0) return false;
    hipLaunchKernelGGL((k_rfft_n<DIR, INV, N, LPB, C, TO, ACC>), dim3((nlines / LPB) * c.g.nz),
                       dim3(RfftShape<N, LPB>::NTHR), 0, c.stream, c.g, tw, in, out);
    return true;
}

// lines per block: along x 4; along y enough lines for 64-byte row segments
template <int DIR, int INV, class C, class TO, int ACC>
static bool rfft_fixed(Ctx& c, const C* tw, const C* in, TO* out, int n, int nlines) {
    constexpr int L = DIR == 0 ? 4 : (sizeof(C) == 4 ? 16 : 8);
    constexpr int L512 = L > 8 ? 8 : 4, L1024 = L > 8 ? 4 : 2;
    return try_rfft_n<DIR, INV, 32, L, C, TO, ACC>(c, tw, in, out, n, nlines) ||
           try_rfft_n<DIR, INV, 64, L, C, TO, ACC>(c, tw, in, out, n, nlines) ||
           try_rfft_n<DIR, INV, 128, L, C, TO, ACC>(c, tw, in, out, n, nlines) ||
           try_rfft_n<DIR, INV, 256, L, C, TO, ACC>(c, tw, in, out, n, nlines) ||
           try_rfft_n<DIR, INV, 512, L512, C, TO, ACC>(c, tw, in, out, n, nlines) ||
           try_rfft_n<DIR, INV, 1024, L1024, C, TO, ACC>(c, tw, in, out, n, nlines);
}

template <int DIR, int INV, class C, class TO, int ACC>
static int rfft_launch(Ctx& c, const C* tw, const C* in, TO* out) {
    const int N = DIR == 0 ? c.g.nx : c.g.ny;
    const int nlines = DIR == 0 ? c.g.ny : c.g.nx;
    if (rfft_fixed<DIR, INV, C, TO, ACC>(c, tw, in, out, N, nlines)) {
        HIP_OK(hipGetLastError());
        return ICLS_OK;
    }
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
    hipLaunchKernelGGL((k_rfft<DIR, INV, C, TO, ACC>), grd, dim3(nthr), shm, c.stream, c.g, N, LPB, TPL, tw, in, out);
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

template <class TS>
static int thomas_launch(Ctx& c, TS* f, double* y) {
    ProfScope ps(c, "k_thomas");
    int bx = c.g.nx < 64 ? c.g.nx : 64;
    hipLaunchKernelGGL(k_thomas<TS>, dim3((c.g.nx + bx - 1) / bx, c.g.ny, 1), dim3(bx), 0, c.stream, c.g,
                       (const double*)c.xyrt, f, c.thomas_d, y);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// Residual of the discrete Poisson problem the transforms and the Thomas solve
// invert exactly (periodic 3-point second differences in x and y, the Neumann
// tridiagonal rows of initsolver.f90:36-49 in z): r = rhs - L p, in double.
template <class T>
__global__ void k_residual(Grid g, const T* __restrict__ rhs, const double* __restrict__ p, T* __restrict__ r) {
    int i0 = blockIdx.x * 64 + threadIdx.x, j0 = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z + 1;
    if (i0 >= g.nx || j0 >= g.ny) return;
    const long q = gidx(g, i0, j0, k);
    const double pc = p[q];
    const double ax = g.dxi * g.dxi, ay = g.dyi * g.dyi, az = g.dzi * g.dzi;
    double lx = (p[gidx(g, wrap(i0 + 1, g.nx), j0, k)] - 2. * pc + p[gidx(g, wrap(i0 - 1, g.nx), j0, k)]) * ax;
    double ly = (p[gidx(g, i0, wrap(j0 + 1, g.ny), k)] - 2. * pc + p[gidx(g, i0, wrap(j0 - 1, g.ny), k)]) * ay;
    double lz;
    if (k == 1) lz = (p[gidx(g, i0, j0, 2)] - pc) * az;
    else if (k == g.nz) lz = (p[gidx(g, i0, j0, k - 1)] - pc) * az;
    else lz = (p[gidx(g, i0, j0, k + 1)] - 2. * pc + p[gidx(g, i0, j0, k - 1)]) * az;
    r[q] = (T)((double)rhs[q] - (lx + ly + lz));
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

// one float-transform solve of L x = src (src float, overwritten as work array),
// result written to (ACC = 0) or added into (ACC = 1) the double pressure
template <int ACC>
static int solve_fp32_pass(Ctx& c, const float* src) {
    const float* twx = as<float>(c.tw_x_f);
    const float* twy = as<float>(c.tw_y_f);
    float* w = as<float>(c.pf);
    int rc;
    {
        ProfScope ps(c, "k_rfft_x_fwd");
        rc = rfft_launch<0, 0, float, float, 0>(c, twx, src, w);
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
    return rfft_launch<0, 1, float, double, ACC>(c, twx, w, c.p);
}

static int solve_mixed(Ctx& c) {
    int rc = solve_fp32_pass<0>(c, as<float>(c.frhs));
    if (rc) return rc;
    {
        ProfScope ps(c, "k_residual");
        dim3 blk(64, 4, 1);
        hipLaunchKernelGGL(k_residual<float>, grid3(c.g, c.g.nz, blk), blk, 0, c.stream, c.g, as<float>(c.frhs),
                           (const double*)c.p, as<float>(c.pf));
        HIP_OK(hipGetLastError());
    }
    return solve_fp32_pass<1>(c, as<float>(c.pf));
}

int launch_solve(Ctx& c) { return c.fp32 ? solve_mixed(c) : solve_fp64(c); }

}  // namespace icls
