// k_poisson.hip -- solver2d (solver.f90:68-196): real FFT along x, real FFT along y,
// Thomas elimination along z per Fourier mode, inverse FFTs.  On one GPU the
// 2DECOMP pencil transposes (solver.f90:82-83, 98, 113, 165, 179, 192-193) are
// identities: every pass works in place on the same i-fastest array.
//
// The transforms follow VRFFTF / VRFFTB semantics (libs/fft/fft.f:1073-1145,
// 374-450): half-complex packing R(1) = Re F0, R(2k) = Re Fk, R(2k+1) = Im Fk,
// R(N) = Re F(N/2), scaled by sqrt(1/N) both ways.  A length-N real line is
// transformed as a length-N/2 complex Stockham FFT in LDS plus the even/odd split.
#include "icls_internal.hpp"

namespace icls {

struct cplx { double re, im; };
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx cconj(cplx a) { return {a.re, -a.im}; }

// In-LDS Stockham radix-2 FFT of M points for one line; TPL threads per line.
// tw[t] = exp(+2 pi i t / N) for t < N/2 (N = 2M); sign = -1 forward, +1 inverse.
__device__ __forceinline__ cplx* stockham(cplx* x, cplx* y, int M, int lane, int TPL, const cplx* tw, int sign) {
    for (int Ns = 1; Ns < M; Ns <<= 1) {
        for (int j = lane; j < M / 2; j += TPL) {
            int kk = j & (Ns - 1);
            cplx a = x[j], b = x[j + M / 2];
            cplx w = tw[kk * (M / Ns)];      // exp(+2 pi i kk / (2 Ns))
            if (sign < 0) w.im = -w.im;
            b = cmul(b, w);
            int d = ((j - kk) << 1) + kk;
            y[d] = cadd(a, b);
            y[d + Ns] = csub(a, b);
        }
        __syncthreads();
        cplx* t = x; x = y; y = t;
    }
    return x;
}

// DIR 0: lines along x (contiguous), blocks of LPB consecutive j at one k.
// DIR 1: lines along y (stride nx), blocks of LPB consecutive i at one k.
template <int DIR, int INV>
__global__ void k_rfft(Grid g, int N, int LPB, int TPL, const double* __restrict__ twg, double* __restrict__ f) {
    extern __shared__ cplx smem[];
    const int M = N / 2;
    cplx* bufa = smem;
    cplx* bufb = smem + (long)LPB * M;
    cplx* tw = smem + 2L * LPB * M;            // N/2 twiddles
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int k = blockIdx.y + 1;
    const int line0 = blockIdx.x * LPB;
    for (int t = tid; t < M; t += nthr) tw[t] = {twg[2 * t], twg[2 * t + 1]};
    const double scale = sqrt(1. / (double)N);
    // ---- load
    if (!INV) {
        for (int e = tid; e < LPB * N; e += nthr) {
            int ln, n;
            if (DIR == 0) { ln = e / N; n = e - ln * N; } else { n = e / LPB; ln = e - n * LPB; }
            long q = DIR == 0 ? gidx(g, n, line0 + ln, k) : gidx(g, line0 + ln, n, k);
            double v = f[q];
            cplx* z = bufa + (long)ln * M + (n >> 1);
            if (n & 1) z->im = v; else z->re = v;
        }
    } else {
        // half-complex -> X_k, then Z_k = (X_k + conj X_{M-k}) + i w^k (X_k - conj X_{M-k})
        for (int e = tid; e < LPB * N; e += nthr) {
            int ln, n;
            if (DIR == 0) { ln = e / N; n = e - ln * N; } else { n = e / LPB; ln = e - n * LPB; }
            long q = DIR == 0 ? gidx(g, n, line0 + ln, k) : gidx(g, line0 + ln, n, k);
            double v = f[q];
            // stage the raw half-complex values in bufb (as doubles)
            ((double*)(bufb + (long)ln * M))[n] = v;
        }
        __syncthreads();
        for (int e = tid; e < LPB * M; e += nthr) {
            int ln = e / M, kk = e - ln * M;
            const double* r = (const double*)(bufb + (long)ln * M);
            cplx X, Y;   // X_k and X_{M-k}
            if (kk == 0) { X = {r[0], 0.}; Y = {r[N - 1], 0.}; }
            else { X = {r[2 * kk - 1], r[2 * kk]}; Y = {r[2 * (M - kk) - 1], r[2 * (M - kk)]}; }
            cplx cy = cconj(Y);
            cplx s = cadd(X, cy), dlt = csub(X, cy);
            cplx w = tw[kk];
            cplx t = cmul(w, dlt);          // w^k (X_k - conj X_{M-k})
            cplx it = {-t.im, t.re};        // i * t
            bufa[(long)ln * M + kk] = cadd(s, it);
        }
    }
    __syncthreads();
    // ---- transform each line
    const int ln = tid / TPL, lane = tid - ln * TPL;
    cplx* res = nullptr;
    if (ln < LPB) res = stockham(bufa + (long)ln * M, bufb + (long)ln * M, M, lane, TPL, tw, INV ? 1 : -1);
    else {
        // threads beyond LPB*TPL still take part in the barriers
        for (int Ns = 1; Ns < M; Ns <<= 1) __syncthreads();
    }
    // every line ends in the same buffer
    int nst = 0;
    for (int Ns = 1; Ns < M; Ns <<= 1) nst++;
    cplx* outb = (nst & 1) ? bufb : bufa;
    (void)res;
    // ---- store
    if (!INV) {
        // X_k = (Z_k + conj Z_{M-k})/2 + w^-k (Z_k - conj Z_{M-k})/(2i), k = 0..M
        for (int e = tid; e < LPB * N; e += nthr) {
            int lnn, n;
            if (DIR == 0) { lnn = e / N; n = e - lnn * N; } else { n = e / LPB; lnn = e - n * LPB; }
            const cplx* Z = outb + (long)lnn * M;
            double v;
            if (n == 0) v = Z[0].re + Z[0].im;
            else if (n == N - 1) v = Z[0].re - Z[0].im;
            else {
                int kk = (n + 1) >> 1;
                cplx a = Z[kk], b = cconj(Z[M - kk]);
                cplx ev = {(a.re + b.re) * 0.5, (a.im + b.im) * 0.5};
                cplx od = {(a.im - b.im) * 0.5, -(a.re - b.re) * 0.5};   // (a - b)/(2i)
                cplx w = tw[kk];
                w.im = -w.im;
                cplx X = cadd(ev, cmul(w, od));
                v = (n & 1) ? X.re : X.im;
            }
            long q = DIR == 0 ? gidx(g, n, line0 + lnn, k) : gidx(g, line0 + lnn, n, k);
            f[q] = scale * v;
        }
    } else {
        for (int e = tid; e < LPB * N; e += nthr) {
            int lnn, n;
            if (DIR == 0) { lnn = e / N; n = e - lnn * N; } else { n = e / LPB; lnn = e - n * LPB; }
            cplx z = outb[(long)lnn * M + (n >> 1)];
            double v = (n & 1) ? z.im : z.re;
            long q = DIR == 0 ? gidx(g, n, line0 + lnn, k) : gidx(g, line0 + lnn, n, k);
            f[q] = scale * v;
        }
    }
}

// The same transform with the line length N and the lines per block LPB fixed at
// compile time: every index split is a shift and every stage loop unrolls.
template <int DIR, int INV, int N, int LPB>
__global__ void __launch_bounds__(LPB*((N / 4) < 64 ? (N / 4) : 64) < 64 ? 64 : LPB*((N / 4) < 64 ? (N / 4) : 64))
k_rfft_n(Grid g, const double* __restrict__ twg, double* __restrict__ f) {
    constexpr int M = N / 2;
    constexpr int TPL = (M / 2) < 64 ? (M / 2) : 64;
    constexpr int NTHR = LPB * TPL < 64 ? 64 : LPB * TPL;
    __shared__ cplx bufa[LPB * M];
    __shared__ cplx bufb[LPB * M];
    __shared__ cplx tw[M];
    const int tid = threadIdx.x;
    const int k = blockIdx.y + 1;
    const int line0 = blockIdx.x * LPB;
    for (int t = tid; t < M; t += NTHR) tw[t] = {twg[2 * t], twg[2 * t + 1]};
    const double scale = sqrt(1. / (double)N);
    auto gq = [&](int ln, int n) -> long { return DIR == 0 ? gidx(g, n, line0 + ln, k) : gidx(g, line0 + ln, n, k); };
    auto split = [&](int e, int& ln, int& n) {
        if (DIR == 0) { ln = e / N; n = e % N; } else { n = e / LPB; ln = e % LPB; }
    };
    if (!INV) {
#pragma unroll 4
        for (int e = tid; e < LPB * N; e += NTHR) {
            int ln, n;
            split(e, ln, n);
            double v = f[gq(ln, n)];
            cplx* z = bufa + ln * M + (n >> 1);
            if (n & 1) z->im = v; else z->re = v;
        }
    } else {
#pragma unroll 4
        for (int e = tid; e < LPB * N; e += NTHR) {
            int ln, n;
            split(e, ln, n);
            ((double*)(bufb + ln * M))[n] = f[gq(ln, n)];
        }
        __syncthreads();
        for (int e = tid; e < LPB * M; e += NTHR) {
            int ln = e / M, kk = e % M;
            const double* r = (const double*)(bufb + ln * M);
            cplx X, Y;
            if (kk == 0) { X = {r[0], 0.}; Y = {r[N - 1], 0.}; }
            else { X = {r[2 * kk - 1], r[2 * kk]}; Y = {r[2 * (M - kk) - 1], r[2 * (M - kk)]}; }
            cplx cy = cconj(Y);
            cplx s = cadd(X, cy), dlt = csub(X, cy);
            cplx t = cmul(tw[kk], dlt);
            cplx it = {-t.im, t.re};
            bufa[ln * M + kk] = cadd(s, it);
        }
    }
    __syncthreads();
    const int ln = tid / TPL, lane = tid % TPL;
    cplx* x = bufa + (ln < LPB ? ln : 0) * M;
    cplx* y = bufb + (ln < LPB ? ln : 0) * M;
#pragma unroll
    for (int Ns = 1; Ns < M; Ns <<= 1) {
        if (ln < LPB) {
#pragma unroll
            for (int j = lane; j < M / 2; j += TPL) {
                int kk = j & (Ns - 1);
                cplx a = x[j], b = x[j + M / 2];
                cplx w = tw[kk * (M / Ns)];
                if (!INV) w.im = -w.im;
                b = cmul(b, w);
                int d = ((j - kk) << 1) + kk;
                y[d] = cadd(a, b);
                y[d + Ns] = csub(a, b);
            }
        }
        __syncthreads();
        cplx* t = x; x = y; y = t;
    }
    constexpr int NST = (M >= 2) ? __builtin_ctz(M) : 0;
    cplx* outb = (NST & 1) ? bufb : bufa;
    if (!INV) {
#pragma unroll 4
        for (int e = tid; e < LPB * N; e += NTHR) {
            int lnn, n;
            split(e, lnn, n);
            const cplx* Z = outb + lnn * M;
            double v;
            if (n == 0) v = Z[0].re + Z[0].im;
            else if (n == N - 1) v = Z[0].re - Z[0].im;
            else {
                int kk = (n + 1) >> 1;
                cplx a = Z[kk], b = cconj(Z[M - kk]);
                cplx ev = {(a.re + b.re) * 0.5, (a.im + b.im) * 0.5};
                cplx od = {(a.im - b.im) * 0.5, -(a.re - b.re) * 0.5};
                cplx w = tw[kk];
                w.im = -w.im;
                cplx X = cadd(ev, cmul(w, od));
                v = (n & 1) ? X.re : X.im;
            }
            f[gq(lnn, n)] = scale * v;
        }
    } else {
#pragma unroll 4
        for (int e = tid; e < LPB * N; e += NTHR) {
            int lnn, n;
            split(e, lnn, n);
            cplx z = outb[lnn * M + (n >> 1)];
            f[gq(lnn, n)] = scale * ((n & 1) ? z.im : z.re);
        }
    }
}

template <int DIR, int INV, int N, int LPB>
static bool try_rfft_n(Ctx& c, double* f, int n, int nlines) {
    if (n != N || nlines % LPB != 0) return false;
    constexpr int TPL = (N / 4) < 64 ? (N / 4) : 64;
    constexpr int NTHR = LPB * TPL < 64 ? 64 : LPB * TPL;
    hipLaunchKernelGGL((k_rfft_n<DIR, INV, N, LPB>), dim3(nlines / LPB, c.g.nz, 1), dim3(NTHR), 0, c.stream, c.g,
                       (const double*)(DIR == 0 ? c.tw_x : c.tw_y), f);
    return true;
}

template <int DIR, int INV>
static bool rfft_fixed(Ctx& c, double* f, int n, int nlines) {
    constexpr int L = DIR == 0 ? 4 : 8;
    return try_rfft_n<DIR, INV, 32, L>(c, f, n, nlines) || try_rfft_n<DIR, INV, 64, L>(c, f, n, nlines) ||
           try_rfft_n<DIR, INV, 128, L>(c, f, n, nlines) || try_rfft_n<DIR, INV, 256, L>(c, f, n, nlines) ||
           try_rfft_n<DIR, INV, 512, 4>(c, f, n, nlines) || try_rfft_n<DIR, INV, 1024, 2>(c, f, n, nlines);
}

static int rfft_launch(Ctx& c, double* f, int dir, int inverse) {
    const int N = dir == 0 ? c.g.nx : c.g.ny;
    const int nlines = dir == 0 ? c.g.ny : c.g.nx;
    bool done;
    if (dir == 0) done = inverse ? rfft_fixed<0, 1>(c, f, N, nlines) : rfft_fixed<0, 0>(c, f, N, nlines);
    else done = inverse ? rfft_fixed<1, 1>(c, f, N, nlines) : rfft_fixed<1, 0>(c, f, N, nlines);
    if (done) {
        HIP_OK(hipGetLastError());
        return ICLS_OK;
    }
    // any other power of two: the runtime-sized kernel
    const int M = N / 2;
    int LPB = dir == 0 ? 4 : 16;
    if (LPB > nlines) LPB = nlines;
    while (nlines % LPB) LPB >>= 1;
    while (LPB > 1 && (size_t)(2L * LPB * M + M) * sizeof(cplx) > 64 * 1024) LPB >>= 1;
    int TPL = M / 2 < 64 ? M / 2 : 64;
    if (TPL < 1) TPL = 1;
    while (LPB * TPL > 1024) TPL >>= 1;
    int nthr = LPB * TPL;
    if (nthr < 64) nthr = 64;
    size_t shm = (size_t)(2L * LPB * M + M) * sizeof(cplx);
    dim3 grd(nlines / LPB, c.g.nz, 1);
    const double* tw = dir == 0 ? c.tw_x : c.tw_y;
#define RF(D, I) hipLaunchKernelGGL((k_rfft<D, I>), grd, dim3(nthr), shm, c.stream, c.g, N, LPB, TPL, tw, f)
    if (dir == 0 && !inverse) RF(0, 0);
    else if (dir == 0) RF(0, 1);
    else if (!inverse) RF(1, 0);
    else RF(1, 1);
#undef RF
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

int launch_fft_x(Ctx& c, double* f, int inverse) {
    ProfScope ps(c, inverse ? "k_rfft_x_inv" : "k_rfft_x_fwd");
    return rfft_launch(c, f, 0, inverse);
}
int launch_fft_y(Ctx& c, double* f, int inverse) {
    ProfScope ps(c, inverse ? "k_rfft_y_inv" : "k_rfft_y_fwd");
    return rfft_launch(c, f, 1, inverse);
}

// Thomas elimination per (i,j) column (solver.f90:119-160) with the Periodic-branch
// matrix of initsolver.f90:36-49: a = c = dzi^2, b = -(a+c), b(1) += a, b(kmax) += c,
// a(1) = c(kmax) = 0, diagonal shifted by xyrt(i,j).  A zero last pivot (the (0,0)
// mode) gives p(kmax) = 0 as in the reference.
__global__ void k_thomas(Grid g, const double* __restrict__ xyrt, double* __restrict__ p, double* __restrict__ d) {
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
    // fetches a chunk of CH values ahead of the arithmetic that uses them.
    constexpr int CH = 16;
    double buf[CH], dbuf[CH];
    double dprev = 0., pprev = 0.;
    for (int kb = 1; kb <= kmax; kb += CH) {
#pragma unroll
        for (int t = 0; t < CH; t++)
            if (kb + t <= kmax) buf[t] = p[q1 + (long)(kb + t - 1) * sz];
#pragma unroll
        for (int t = 0; t < CH; t++) {
            const int k = kb + t;
            if (k > kmax) break;
            double z;
            if (k == 1) {
                z = 1. / (b1 + lam);
                dprev = cc * z;
                pprev = buf[t] * z;
            } else if (k < kmax) {
                double bb = bmid + lam;
                z = 1. / (bb - a * dprev);
                dprev = cc * z;
                pprev = (buf[t] - a * pprev) * z;
            } else {
                double bb = bN + lam;
                z = bb - a * dprev;
                if (z != 0.) pprev = (buf[t] - a * pprev) / z;
                else pprev = 0.;
            }
            buf[t] = pprev;
            dbuf[t] = dprev;
        }
#pragma unroll
        for (int t = 0; t < CH; t++) {
            const int k = kb + t;
            if (k > kmax) break;
            p[q1 + (long)(k - 1) * sz] = buf[t];
            if (k < kmax) d[q1 + (long)(k - 1) * sz] = dbuf[t];
        }
    }
    // back substitution from k = kmax-1 down to 1; pprev holds p(kmax)
    for (int kt = kmax - 1; kt >= 1; kt -= CH) {
#pragma unroll
        for (int t = 0; t < CH; t++)
            if (kt - t >= 1) {
                long q = q1 + (long)(kt - t - 1) * sz;
                buf[t] = p[q];
                dbuf[t] = d[q];
            }
#pragma unroll
        for (int t = 0; t < CH; t++) {
            if (kt - t < 1) break;
            pprev = buf[t] - dbuf[t] * pprev;
            buf[t] = pprev;
        }
#pragma unroll
        for (int t = 0; t < CH; t++)
            if (kt - t >= 1) p[q1 + (long)(kt - t - 1) * sz] = buf[t];
    }
}

int launch_thomas(Ctx& c, double* f) {
    ProfScope ps(c, "k_thomas");
    int bx = c.g.nx < 64 ? c.g.nx : 64;
    hipLaunchKernelGGL(k_thomas, dim3((c.g.nx + bx - 1) / bx, c.g.ny, 1), dim3(bx), 0, c.stream, c.g,
                       (const double*)c.xyrt, f, c.thomas_d);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

}  // namespace icls
