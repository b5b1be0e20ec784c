// k_mlp.hip -- the learned curvature closure: a 30 -> 64 -> 64 -> 1 tanh MLP with
// BF16 matrix-core (MFMA) forward and backward kernels, Adam on FP32 master weights,
// a training-sample generator that applies the solver's own curvature operator on a
// 4x finer grid, and the inference pass that replaces the finite-difference curvature
// on under-resolved interface cells.
//
// The closure has no counterpart in the reference; it stands in for get_curvature
// (interface.f90:20-119) where the 27-point fit under-resolves the interface.
//
// Matrix products use v_mfma_f32_16x16x16_bf16 with one wavefront per group of 16
// samples.  Operand layouts (lane l, r = 0..3):
//   A (16 x 16): lane l holds A[l % 16][4 (l / 16) + r]
//   B (16 x 16): lane l holds B[4 (l / 16) + r][l % 16]
//   D (16 x 16): lane l holds D[4 (l / 16) + r][l % 16]
// Layer outputs come out in the D layout: lane l holds the activations of samples
// 4 (l / 16) + r for hidden unit 16 nt + l % 16 of tile nt.  That is exactly the A
// layout of the transposed activation, so the weight gradients take the layer
// outputs straight from registers.
#include "icls_internal.hpp"
#include "icls_stencil.hpp"

namespace icls {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = (short)f2bf(a); r[1] = (short)f2bf(b); r[2] = (short)f2bf(c); r[3] = (short)f2bf(d);
    return r;
}
__device__ __forceinline__ bf16x4 ld4(const uint16_t* p) { return *reinterpret_cast<const bf16x4*>(p); }
__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// sum over the 16 lanes that share l / 16 (one sample row of the D layout)
__device__ __forceinline__ float sum16(float v) {
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    return v;
}
// sum over the four lane groups (the sample rows) at a fixed l % 16
__device__ __forceinline__ float sum4g(float v) {
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

struct Fwd {
    f32x4 h1[4], h2[4];   // D layout, tiles of 16 hidden units
    float y[4];           // outputs of samples 4 (l / 16) + r
};

// Forward pass of 16 samples held in LDS (xs: 16 x 32 bf16).  hs: 16 x 64 bf16 LDS
// scratch.  The caller's block is one wavefront.
__device__ __forceinline__ void mlp_fwd16(const uint16_t* xs, uint16_t* hs, const uint16_t* __restrict__ w1b,
                                          const uint16_t* __restrict__ w2b, const float* __restrict__ w, Fwd& o) {
    const int l = threadIdx.x, lr = l & 15, lg = l >> 4;
    const bf16x4 xa0 = ld4(xs + lr * MLP_NPAD + 4 * lg);
    const bf16x4 xa1 = ld4(xs + lr * MLP_NPAD + 16 + 4 * lg);
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        const uint16_t* wr = w1b + (16 * nt + lr) * MLP_NPAD + 4 * lg;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma16(xa0, ld4(wr), acc);
        acc = mfma16(xa1, ld4(wr + 16), acc);
        const float b = w[MLP_OB1 + 16 * nt + lr];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            o.h1[nt][r] = tanhf(acc[r] + b);
            hs[(4 * lg + r) * MLP_NH + 16 * nt + lr] = f2bf(o.h1[nt][r]);
        }
    }
    __syncthreads();
    bf16x4 ha[4];
#pragma unroll
    for (int kt = 0; kt < 4; kt++) ha[kt] = ld4(hs + lr * MLP_NH + 16 * kt + 4 * lg);
    float s[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        const uint16_t* wr = w2b + (16 * nt + lr) * MLP_NH + 4 * lg;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; kt++) acc = mfma16(ha[kt], ld4(wr + 16 * kt), acc);
        const float b = w[MLP_OB2 + 16 * nt + lr], w3 = w[MLP_OW3 + 16 * nt + lr];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            o.h2[nt][r] = tanhf(acc[r] + b);
            s[r] += o.h2[nt][r] * w3;
        }
    }
    const float b3 = w[MLP_OB3];
#pragma unroll
    for (int r = 0; r < 4; r++) o.y[r] = sum16(s[r]) + b3;
    __syncthreads();   // hs may be rewritten by the caller
}

// ---------------------------------------------------------------- batch forward
// y[n] for n rows of x (bf16, 32 per row); n is a multiple of 16.
__global__ void __launch_bounds__(64)
k_mlp_forward(const uint16_t* __restrict__ x, int n, const uint16_t* __restrict__ w1b,
              const uint16_t* __restrict__ w2b, const float* __restrict__ w, float* __restrict__ y) {
    __shared__ __align__(16) uint16_t xs[16 * MLP_NPAD];
    __shared__ __align__(16) uint16_t hs[16 * MLP_NH];
    const int l = threadIdx.x, lr = l & 15, lg = l >> 4;
    for (int grp = blockIdx.x; grp * 16 < n; grp += gridDim.x) {
        const uint16_t* xg = x + (long)grp * 16 * MLP_NPAD;
        // 512 bf16 = 64 lanes x 8
        *reinterpret_cast<bf16x4*>(xs + 8 * l) = ld4(xg + 8 * l);
        *reinterpret_cast<bf16x4*>(xs + 8 * l + 4) = ld4(xg + 8 * l + 4);
        __syncthreads();
        Fwd o;
        mlp_fwd16(xs, hs, w1b, w2b, w, o);
        if (lr == 0)
#pragma unroll
            for (int r = 0; r < 4; r++) y[(long)grp * 16 + 4 * lg + r] = o.y[r];
    }
}

// ---------------------------------------------------------------- gradients
// Loss = mean over the n rows of (y - t)^2.  Block b folds the row groups
// b, b + gridDim.x, ... and writes its partial gradient (MLP_NPARAM values, then the
// partial loss) to part[b].  The bf16 rounding points are those of the oracle:
// the inputs, W1, W2, h1, and the back-propagated dz2 and dz1.
__global__ void __launch_bounds__(64)
k_mlp_grad(const uint16_t* __restrict__ x, const float* __restrict__ t, int n, float scale,
           const uint16_t* __restrict__ w1b, const uint16_t* __restrict__ w2b, const uint16_t* __restrict__ w2tb,
           const float* __restrict__ w, float* __restrict__ part) {
    __shared__ __align__(16) uint16_t xs[16 * MLP_NPAD];
    __shared__ __align__(16) uint16_t hs[16 * MLP_NH];
    const int l = threadIdx.x, lr = l & 15, lg = l >> 4;
    f32x4 gW2[4][4], gW1[4][2];
    float gb1[4], gb2[4], gw3[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
#pragma unroll
        for (int b = 0; b < 4; b++) gW2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
        gW1[a][0] = f32x4{0.f, 0.f, 0.f, 0.f};
        gW1[a][1] = f32x4{0.f, 0.f, 0.f, 0.f};
        gb1[a] = gb2[a] = gw3[a] = 0.f;
    }
    float gb3 = 0.f, loss = 0.f;
    float w3c[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) w3c[nt] = w[MLP_OW3 + 16 * nt + lr];

    for (int grp = blockIdx.x; grp * 16 < n; grp += gridDim.x) {
        const uint16_t* xg = x + (long)grp * 16 * MLP_NPAD;
        *reinterpret_cast<bf16x4*>(xs + 8 * l) = ld4(xg + 8 * l);
        *reinterpret_cast<bf16x4*>(xs + 8 * l + 4) = ld4(xg + 8 * l + 4);
        __syncthreads();
        Fwd o;
        mlp_fwd16(xs, hs, w1b, w2b, w, o);
        float dy[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float e = o.y[r] - t[(long)grp * 16 + 4 * lg + r];
            dy[r] = scale * e;
            if (lr == 0) { loss += e * e; gb3 += dy[r]; }
        }
        // output layer and dz2 = dy w3 (1 - h2^2)
        f32x4 dz2[4];
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                gw3[nt] += o.h2[nt][r] * dy[r];
                dz2[nt][r] = dy[r] * w3c[nt] * (1.f - o.h2[nt][r] * o.h2[nt][r]);
                gb2[nt] += dz2[nt][r];
                hs[(4 * lg + r) * MLP_NH + 16 * nt + lr] = f2bf(dz2[nt][r]);
            }
        // gW2[n][k] += sum_rows dz2[row][n] h1[row][k]
        bf16x4 h1b[4], dzb[4];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            h1b[nt] = pack4(o.h1[nt][0], o.h1[nt][1], o.h1[nt][2], o.h1[nt][3]);
            dzb[nt] = pack4(dz2[nt][0], dz2[nt][1], dz2[nt][2], dz2[nt][3]);
        }
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int nt = 0; nt < 4; nt++) gW2[mt][nt] = mfma16(dzb[mt], h1b[nt], gW2[mt][nt]);
        __syncthreads();
        // dh1 = dz2 W2 (A = dz2 in row layout from LDS, B = W2 via its transpose)
        bf16x4 da[4];
#pragma unroll
        for (int kt = 0; kt < 4; kt++) da[kt] = ld4(hs + lr * MLP_NH + 16 * kt + 4 * lg);
        f32x4 dz1[4];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const uint16_t* wr = w2tb + (16 * nt + lr) * MLP_NH + 4 * lg;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; kt++) acc = mfma16(da[kt], ld4(wr + 16 * kt), acc);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                dz1[nt][r] = acc[r] * (1.f - o.h1[nt][r] * o.h1[nt][r]);
                gb1[nt] += dz1[nt][r];
            }
        }
        // gW1[n][k] += sum_rows dz1[row][n] x[row][k]
        bf16x4 xb[2];
#pragma unroll
        for (int kt = 0; kt < 2; kt++) {
            bf16x4 v;
#pragma unroll
            for (int r = 0; r < 4; r++) v[r] = (short)xs[(4 * lg + r) * MLP_NPAD + 16 * kt + lr];
            xb[kt] = v;
        }
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            const bf16x4 a = pack4(dz1[mt][0], dz1[mt][1], dz1[mt][2], dz1[mt][3]);
            gW1[mt][0] = mfma16(a, xb[0], gW1[mt][0]);
            gW1[mt][1] = mfma16(a, xb[1], gW1[mt][1]);
        }
        __syncthreads();   // xs and hs are rewritten by the next group
    }
    float* out = part + (long)blockIdx.x * (MLP_NPARAM + 1);
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int nrow = 16 * mt + 4 * lg + r;
#pragma unroll
            for (int nt = 0; nt < 4; nt++) out[MLP_OW2 + nrow * MLP_NH + 16 * nt + lr] = gW2[mt][nt][r];
            out[MLP_OW1 + nrow * MLP_NPAD + lr] = gW1[mt][0][r];
            out[MLP_OW1 + nrow * MLP_NPAD + 16 + lr] = gW1[mt][1][r];
        }
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        const float a = sum4g(gb1[nt]), b = sum4g(gb2[nt]), c = sum4g(gw3[nt]);
        if (lg == 0) {
            out[MLP_OB1 + 16 * nt + lr] = a;
            out[MLP_OB2 + 16 * nt + lr] = b;
            out[MLP_OW3 + 16 * nt + lr] = c;
        }
    }
    gb3 = sum4g(gb3);
    loss = sum4g(loss);
    if (l == 0) {
        out[MLP_OB3] = gb3;
        out[MLP_NPARAM] = loss;
    }
}

// Partials -> gradient (fixed block order), mean loss into grad[MLP_NPARAM]; with
// lr > 0 also one Adam step on the master weights and their bf16 copies.
__global__ void k_mlp_reduce_adam(const float* __restrict__ part, int nblk, float inv_n, float* __restrict__ grad,
                                  float* __restrict__ w, float* __restrict__ m, float* __restrict__ v, float lr,
                                  float bc1, float bc2, uint16_t* __restrict__ w1b, uint16_t* __restrict__ w2b,
                                  uint16_t* __restrict__ w2tb) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > MLP_NPARAM) return;
    float g = 0.f;
    for (int b = 0; b < nblk; b++) g += part[(long)b * (MLP_NPARAM + 1) + p];
    if (p == MLP_NPARAM) { grad[p] = g * inv_n; return; }
    grad[p] = g;
    if (lr <= 0.f) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float wn = w[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    w[p] = wn;
    if (p < MLP_OB1) {
        w1b[p] = f2bf(wn);
    } else if (p >= MLP_OW2 && p < MLP_OB2) {
        const int q = p - MLP_OW2, nrow = q / MLP_NH, kcol = q % MLP_NH;
        w2b[q] = f2bf(wn);
        w2tb[kcol * MLP_NH + nrow] = f2bf(wn);
    }
}

// bf16 copies of freshly set master weights
__global__ void k_mlp_cast(const float* __restrict__ w, uint16_t* __restrict__ w1b, uint16_t* __restrict__ w2b,
                           uint16_t* __restrict__ w2tb) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < MLP_OB1) w1b[p] = f2bf(w[p]);
    if (p >= MLP_OW2 && p < MLP_OB2) {
        const int q = p - MLP_OW2, nrow = q / MLP_NH, kcol = q % MLP_NH;
        w2b[q] = f2bf(w[p]);
        w2tb[kcol * MLP_NH + nrow] = f2bf(w[p]);
    }
}

// ---------------------------------------------------------------- training samples
__device__ __forceinline__ unsigned long long splitmix(unsigned long long& s) {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double urand(unsigned long long& s) { return (double)(splitmix(s) >> 11) * 0x1.0p-53; }

// Sample i: a sphere (drop, phi = |x - c| - R) or its complement (bubble,
// phi = R - |x - c|) of radius R in [r_min, r_max] dz, placed so that the coarse cell
// centred at the origin lies within 4 dz of the interface.  Features: the exact
// signed distance on the coarse 3 x 3 x 3 stencil over dz and the unit normal of the
// least-squares gradient.  Reference curvature: the mean of the solver's curvature
// operator (curv_lsq) on the 8 cells of a 4x finer grid around the origin.  The
// network learns the correction to the coarse operator's value, so the target is
// (kappa_fine - kappa_coarse) dz, with kappa_fine limited to the +-1/dz the solver's
// curvature is limited to; fd (optional) receives kappa_coarse dz.
//
// With gate > 0 a sample is drawn again (up to 64 times) until the coarse cell is one
// the closure is consulted on: |grad phi| >= 0.9 in the least-squares fit and
// |kappa_coarse| dz >= gate.
// The closure's inputs from the 27-point stencil v (x fastest, offsets -1 .. 1 in each
// direction) and the stencil's least-squares gradient (gx, gy, gz):
//   f[q], q != 13  (v[q] - v[13] - g . r_q) / dz, the stencil less the value and the slope
//                  at the centre, which leaves its quadratic part: numbers of the size of
//                  kappa dz that bfloat16 holds to three digits (the raw |v| / dz, up to 6,
//                  would lose the second differences the curvature lives in to rounding);
//   f[13]          v[13] / dz, where the interface lies;
//   f[27..29]      the unit normal g / |g|.
__device__ inline void mlp_features(const double* v, double gx, double gy, double gz, double dx, double dy, double dz,
                                    float* f) {
    int q = 0;
    for (int c = -1; c <= 1; c++)
        for (int b = -1; b <= 1; b++)
            for (int a = -1; a <= 1; a++, q++)
                f[q] = (float)((v[q] - v[13] - (gx * (a * dx) + gy * (b * dy) + gz * (c * dz))) / dz);
    f[13] = (float)(v[13] / dz);
    const double gn = sqrt(gx * gx + gy * gy + gz * gz);
    f[27] = (float)(gx / gn);
    f[28] = (float)(gy / gn);
    f[29] = (float)(gz / gn);
    f[30] = 0.f;
    f[31] = 0.f;
}

__global__ void k_mlp_samples(long n, unsigned long long seed, double r_min, double r_max, double gate, double dx,
                              double dy, double dz, const double* __restrict__ A, const double* __restrict__ Af,
                              uint16_t* __restrict__ x, float* __restrict__ t, float* __restrict__ fd) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long s = seed ^ (0xD1B54A32D192ED03ull * (unsigned long long)(i + 1));
    double R, sg, cx, cy, cz, kc, v[27];
    for (int attempt = 0;; attempt++) {
        R = (r_min + (r_max - r_min) * urand(s)) * dz;
        const double zc = 2. * urand(s) - 1., ang = 2. * 3.14159265358979323846 * urand(s);
        const double rs = sqrt(fmax(0., 1. - zc * zc));
        const double nx = rs * cos(ang), ny = rs * sin(ang), nz = zc;
        sg = urand(s) < 0.5 ? 1. : -1.;
        // a cell of the re_curvature shell |phi| < 2 dz, with half a cell of margin
        const double dlo = fmax(0.1 * R, R - 2.5 * dz), dhi = R + 2.5 * dz;
        const double dist = dlo + (dhi - dlo) * urand(s);
        cx = -dist * nx;
        cy = -dist * ny;
        cz = -dist * nz;
        int q = 0;
        for (int c = -1; c <= 1; c++)
            for (int b = -1; b <= 1; b++)
                for (int a = -1; a <= 1; a++) {
                    const double ex = a * dx - cx, ey = b * dy - cy, ez = c * dz - cz;
                    v[q++] = sg * (sqrt(ex * ex + ey * ey + ez * ez) - R);
                }
        double mc;
        kc = curv_lsq<double>(A, v, 1. / dz, mc);
        if (gate <= 0. || attempt >= 63 || (mc >= 0.9 && fabs(kc) * dz >= gate)) break;
    }
    auto phi = [&](double px, double py, double pz) {
        const double ex = px - cx, ey = py - cy, ez = pz - cz;
        return sg * (sqrt(ex * ex + ey * ey + ez * ez) - R);
    };
    double gx = 0., gy = 0., gz = 0.;
    for (int k = 0; k < 27; k++) {
        gx += A[1 + 10 * k] * v[k];
        gy += A[2 + 10 * k] * v[k];
        gz += A[3 + 10 * k] * v[k];
    }
    float fe[MLP_NPAD];
    mlp_features(v, gx, gy, gz, dx, dy, dz, fe);
    uint16_t* xr = x + i * MLP_NPAD;
    for (int k = 0; k < MLP_NPAD; k++) xr[k] = f2bf(fe[k]);
    // the fine grid: spacing h/4, cells centred at +-h/8 around the origin
    const double fx = dx / 4., fy = dy / 4., fz = dz / 4.;
    double ksum = 0.;
    for (int oc = 0; oc < 8; oc++) {
        const double ox = (oc & 1 ? 0.5 : -0.5) * fx, oy = (oc & 2 ? 0.5 : -0.5) * fy,
                     oz = (oc & 4 ? 0.5 : -0.5) * fz;
        double vf[27];
        int qq = 0;
        for (int c = -1; c <= 1; c++)
            for (int b = -1; b <= 1; b++)
                for (int a = -1; a <= 1; a++) vf[qq++] = phi(ox + a * fx, oy + b * fy, oz + c * fz);
        double m;
        ksum += curv_lsq<double>(Af, vf, 1. / fz, m);
    }
    const double kf = fmin(fmax(ksum / 8., -1. / dz), 1. / dz);
    t[i] = (float)((kf - kc) * dz);
    if (fd) fd[i] = (float)(kc * dz);
}

// ---------------------------------------------------------------- inference
// The listed cells (k_curvature): features from phi; the closure's kappa is the cell's
// own finite-difference value (kfd_in, before re_curvature) plus the network's
// correction, and it replaces the smoothed value in `kappa` when the correction is
// within tol / dz (the smoothed finite-difference value stays otherwise).  The
// smoothing averages the 27 neighbours' curvatures, which for a small drop includes
// the clipped ones near its centre; the closure's value is the curvature of the level
// surface through the cell, which is what the jump interpolation wants.
// BATCH: the cells of a batched curvature launch, whole-grid indices (g is the whole
// grid) with the window of each in list_t, whose level set is wb.ls[window].
template <class T, bool BATCH>
__global__ void __launch_bounds__(64)
k_mlp_curv(Grid g, const int* __restrict__ list, const int* __restrict__ list_t, WinBatch wb,
           const int* __restrict__ count, int cap, const T* __restrict__ A, const T* __restrict__ phi,
           const T* __restrict__ kfd_in, T* __restrict__ kappa, const uint16_t* __restrict__ w1b,
           const uint16_t* __restrict__ w2b, const float* __restrict__ w, float tol, int* __restrict__ total) {
    __shared__ __align__(16) uint16_t xs[16 * MLP_NPAD];
    __shared__ __align__(16) uint16_t hs[16 * MLP_NH];
    __shared__ long qs[16];
    const int l = threadIdx.x, lg = l >> 4, lr = l & 15;
    const int cnt = min(count[0], cap);
    const double dz = g.dz;
    if (blockIdx.x == 0 && l == 0) atomicAdd(total, cnt);   // the step's listed cells, all level sets
    for (int grp = blockIdx.x; grp * 16 < cnt; grp += gridDim.x) {
        // lanes 0..15 build the feature rows of the group's cells
        if (l < 16) {
            const int e = grp * 16 + l;
            long q = -1;
            float f[MLP_NPAD];
#pragma unroll
            for (int k = 0; k < MLP_NPAD; k++) f[k] = 0.f;
            if (e < cnt) {
                const int c = list[e];
                const T* ph = phi;
                if constexpr (BATCH) ph = (const T*)wb.ls[list_t[e]];
                // the list holds i0 + sy j0 + sz (k - 1) (k_curvature): decode with the
                // strides, which differ from nx and ny inside a level-set window
                const int i0 = (int)(c % g.sy), j0 = (int)((c / g.sy) % (g.sz / g.sy)), k = (int)(c / g.sz) + 1;
                q = gidx(g, i0, j0, k);
                double v[27];
                int qq = 0;
                for (int cc = -1; cc <= 1; cc++)
                    for (int b = -1; b <= 1; b++)
                        for (int a = -1; a <= 1; a++)
                            v[qq++] = (double)ph[gidx(g, wrap(i0 + a, g.nx), wrap(j0 + b, g.ny), k + cc)];
                double gx = 0., gy = 0., gz = 0.;
                for (int kk = 0; kk < 27; kk++) {
                    gx += (double)A[1 + 10 * kk] * v[kk];
                    gy += (double)A[2 + 10 * kk] * v[kk];
                    gz += (double)A[3 + 10 * kk] * v[kk];
                }
                mlp_features(v, gx, gy, gz, g.dx, g.dy, dz, f);
                if (!(sqrt(gx * gx + gy * gy + gz * gz) >= 0.9)) q = -1;   // keep the smoothed value
            }
#pragma unroll
            for (int k = 0; k < MLP_NPAD; k++) xs[l * MLP_NPAD + k] = f2bf(f[k]);
            qs[l] = q;
        }
        __syncthreads();
        Fwd o;
        mlp_fwd16(xs, hs, w1b, w2b, w, o);
        if (lr == 0)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const long q = qs[4 * lg + r];
                if (q < 0) continue;
                const double kfd = (double)kfd_in[q];
                double km = kfd + (double)o.y[r] / dz;
                const double kmax = 1. / dz;
                km = km < -kmax ? -kmax : (km > kmax ? kmax : km);
                if (isfinite(km) && fabs(km - kfd) * dz <= (double)tol) kappa[q] = (T)km;
            }
        __syncthreads();
    }
}

__global__ void k_zero_int(int* p) {
    if (threadIdx.x == 0) p[0] = 0;
}
__global__ void k_zero_int2(int* p, int* q) {
    if (threadIdx.x == 0) { p[0] = 0; q[0] = 0; }
}

// ---------------------------------------------------------------- host side
void pseudo_inverse(double dx, double dy, double dz, double* out);

int mlp_alloc(Ctx& c) {
    Mlp& m = c.mlp;
    HIP_OK(hipMalloc((void**)&m.w, sizeof(float) * MLP_NPARAM));
    HIP_OK(hipMalloc((void**)&m.adam_m, sizeof(float) * MLP_NPARAM));
    HIP_OK(hipMalloc((void**)&m.adam_v, sizeof(float) * MLP_NPARAM));
    HIP_OK(hipMemset(m.adam_m, 0, sizeof(float) * MLP_NPARAM));
    HIP_OK(hipMemset(m.adam_v, 0, sizeof(float) * MLP_NPARAM));
    HIP_OK(hipMalloc((void**)&m.w1b, sizeof(uint16_t) * MLP_NH * MLP_NPAD));
    HIP_OK(hipMalloc((void**)&m.w2b, sizeof(uint16_t) * MLP_NH * MLP_NH));
    HIP_OK(hipMalloc((void**)&m.w2tb, sizeof(uint16_t) * MLP_NH * MLP_NH));
    HIP_OK(hipMalloc((void**)&m.gpart, sizeof(float) * MLP_GBLOCKS * (MLP_NPARAM + 1)));
    HIP_OK(hipMalloc((void**)&m.grad, sizeof(float) * (MLP_NPARAM + 1)));
    std::vector<double> af(270);
    pseudo_inverse(c.g.dx / 4., c.g.dy / 4., c.g.dz / 4., af.data());
    HIP_OK(hipMalloc((void**)&m.pinvA_fine, sizeof(double) * 270));
    HIP_OK(hipMemcpy(m.pinvA_fine, af.data(), sizeof(double) * 270, hipMemcpyHostToDevice));
    m.list_cap = (int)std::min<long>(c.g.sz * (long)c.g.nz, 1L << 22);
    HIP_OK(hipMalloc((void**)&m.list, sizeof(int) * m.list_cap));
    HIP_OK(hipMalloc((void**)&m.list_t, sizeof(int) * m.list_cap));
    HIP_OK(hipMalloc((void**)&m.count, sizeof(int)));
    HIP_OK(hipMemset(m.count, 0, sizeof(int)));
    HIP_OK(hipMalloc((void**)&m.total, sizeof(int)));
    HIP_OK(hipMemset(m.total, 0, sizeof(int)));
    // deterministic initial weights: uniform(-1, 1) / sqrt(fan_in), zero biases and
    // zero padding columns
    std::vector<float> w(MLP_NPARAM, 0.f);
    unsigned long long s = 0x1C15ull;
    auto u = [&]() {
        unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return 2. * ((double)(z >> 11) * 0x1.0p-53) - 1.;
    };
    for (int nrow = 0; nrow < MLP_NH; nrow++)
        for (int k = 0; k < MLP_NIN; k++) w[MLP_OW1 + nrow * MLP_NPAD + k] = (float)(u() / sqrt((double)MLP_NIN));
    for (int p = MLP_OW2; p < MLP_OB2; p++) w[p] = (float)(u() / sqrt((double)MLP_NH));
    for (int p = MLP_OW3; p < MLP_OB3; p++) w[p] = (float)(u() / sqrt((double)MLP_NH));
    return mlp_set_weights(c, w.data());
}

void mlp_free(Ctx& c) {
    Mlp& m = c.mlp;
    void* ptrs[] = {m.w, m.adam_m, m.adam_v, m.w1b, m.w2b, m.w2tb, m.gpart, m.grad, m.pinvA_fine, m.list, m.list_t,
                    m.count, m.total};
    for (void* p : ptrs) if (p) hipFree(p);
    m = Mlp();
}

int mlp_set_weights(Ctx& c, const float* w) {
    Mlp& m = c.mlp;
    HIP_OK(hipMemcpyAsync(m.w, w, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_mlp_cast, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, c.stream, (const float*)m.w, m.w1b,
                       m.w2b, m.w2tb);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemsetAsync(m.adam_m, 0, sizeof(float) * MLP_NPARAM, c.stream));
    HIP_OK(hipMemsetAsync(m.adam_v, 0, sizeof(float) * MLP_NPARAM, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    m.adam_t = 0;
    return ICLS_OK;
}

int mlp_forward(Ctx& c, const uint16_t* x, int n, float* y) {
    if (n <= 0 || n % 16) { set_error("forward batch must be a positive multiple of 16 (got %d)", n); return ICLS_ERR_PARAM; }
    ProfScope ps(c, "k_mlp_forward");
    int nb = n / 16 < 1024 ? n / 16 : 1024;
    hipLaunchKernelGGL(k_mlp_forward, dim3(nb), dim3(64), 0, c.stream, x, n, (const uint16_t*)c.mlp.w1b,
                       (const uint16_t*)c.mlp.w2b, (const float*)c.mlp.w, y);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

// gradient of the mean squared error over n rows into c.mlp.grad (and grad_loss if given)
int mlp_grad(Ctx& c, const uint16_t* x, const float* t, int n, float* grad_loss) {
    if (n <= 0 || n % 16) { set_error("gradient batch must be a positive multiple of 16 (got %d)", n); return ICLS_ERR_PARAM; }
    Mlp& m = c.mlp;
    ProfScope ps(c, "k_mlp_grad");
    hipLaunchKernelGGL(k_mlp_grad, dim3(MLP_GBLOCKS), dim3(64), 0, c.stream, x, t, n, 2.f / (float)n,
                       (const uint16_t*)m.w1b, (const uint16_t*)m.w2b, (const uint16_t*)m.w2tb, (const float*)m.w,
                       m.gpart);
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((MLP_NPARAM + 1 + 255) / 256), dim3(256), 0, c.stream,
                       (const float*)m.gpart, MLP_GBLOCKS, 1.f / (float)n, m.grad, m.w, m.adam_m, m.adam_v, 0.f, 1.f,
                       1.f, m.w1b, m.w2b, m.w2tb);
    HIP_OK(hipGetLastError());
    if (grad_loss) HIP_OK(hipMemcpyAsync(grad_loss, m.grad, sizeof(float) * (MLP_NPARAM + 1), hipMemcpyDeviceToDevice,
                                         c.stream));
    return ICLS_OK;
}

// one Adam step from the partials the last k_mlp_grad left (re-reduced in the same
// fixed order)
int mlp_adam(Ctx& c, float lr) {
    Mlp& m = c.mlp;
    m.adam_t++;
    const float bc1 = 1.f - powf(0.9f, (float)m.adam_t), bc2 = 1.f - powf(0.999f, (float)m.adam_t);
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((MLP_NPARAM + 1 + 255) / 256), dim3(256), 0, c.stream,
                       (const float*)m.gpart, MLP_GBLOCKS, 1.f, m.grad, m.w, m.adam_m, m.adam_v, lr, bc1, bc2, m.w1b,
                       m.w2b, m.w2tb);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

int mlp_samples(Ctx& c, long n, unsigned long long seed, float r_min, float r_max, float gate, uint16_t* x,
                float* t, float* fd) {
    ProfScope ps(c, "k_mlp_samples");
    hipLaunchKernelGGL(k_mlp_samples, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, c.stream, n, seed,
                       (double)r_min, (double)r_max, (double)gate, c.g.dx, c.g.dy, c.g.dz, (const double*)c.pinvA,
                       (const double*)c.mlp.pinvA_fine, x, t, fd);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

int launch_mlp_reset_list(Ctx& c) {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c.stream, c.mlp.count);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

template <bool BATCH>
static int mlp_curv_launch(Ctx& c, const WinBatch& wb, const double* phi, const double* kappa_fd, double* kappa_out) {
    ProfScope ps(c, "k_mlp_curv");
    Mlp& m = c.mlp;
    const int nb = 512;
    if (c.fp32)
        hipLaunchKernelGGL((k_mlp_curv<float, BATCH>), dim3(nb), dim3(64), 0, c.stream, c.g, (const int*)m.list,
                           (const int*)m.list_t, wb, (const int*)m.count, m.list_cap, as<float>(c.pinvA_f),
                           as<float>(phi), as<float>(kappa_fd), as<float>(kappa_out), (const uint16_t*)m.w1b,
                           (const uint16_t*)m.w2b, (const float*)m.w, m.tol, m.total);
    else
        hipLaunchKernelGGL((k_mlp_curv<double, BATCH>), dim3(nb), dim3(64), 0, c.stream, c.g, (const int*)m.list,
                           (const int*)m.list_t, wb, (const int*)m.count, m.list_cap, (const double*)c.pinvA, phi,
                           kappa_fd, kappa_out, (const uint16_t*)m.w1b, (const uint16_t*)m.w2b, (const float*)m.w,
                           m.tol, m.total);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}

int launch_mlp_curv(Ctx& c, const double* phi, const double* kappa_fd, double* kappa_out) {
    return mlp_curv_launch<false>(c, WinBatch(), phi, kappa_fd, kappa_out);
}

int launch_mlp_curv_batch(Ctx& c, const WinBatch& wb, const double* kappa_fd, double* kappa_out) {
    if (!c.mlp.list_t) {
        set_error("closure cell list without window entries");
        return ICLS_ERR_STATE;
    }
    return mlp_curv_launch<true>(c, wb, nullptr, kappa_fd, kappa_out);
}

int launch_mlp_reset_list_total(Ctx& c) {
    hipLaunchKernelGGL(k_zero_int2, dim3(1), dim3(64), 0, c.stream, c.mlp.count, c.mlp.total);
    HIP_OK(hipGetLastError());
    return ICLS_OK;
}


}  // namespace icls
