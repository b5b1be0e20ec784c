// icls_stencil.hpp -- pointwise and line-stencil device functions of the ICLS step.
// Each follows the reference expression cited beside it, in the same evaluation
// order; the library is compiled with -ffp-contract=off so no product is fused into
// an add and FP64 results match the reference's operation sequence.
#pragma once
#include <hip/hip_runtime.h>

namespace icls {

// HOUC5 upwind derivative along a line f[-3..3] (hyperPDE.f90:427-433, 441-447, 455-461).
template <typename T>
__device__ __forceinline__ T houc5(T vel, const T* f) {
    if (vel >= T(0))
        return T(1.) / T(60.) * (-T(2.) * f[-3] + T(15.) * f[-2] - T(60.) * f[-1] + T(20.) * f[0] + T(30.) * f[1] - T(3.) * f[2]);
    return T(1.) / T(60.) * (T(2.) * f[3] - T(15.) * f[2] + T(60.) * f[1] - T(20.) * f[0] - T(30.) * f[-1] + T(3.) * f[-2]);
}

// HOUC3 next to the walls, k = 2 and k = kmax-1 (hyperPDE.f90:465-471)
template <typename T>
__device__ __forceinline__ T houc3(T vel, const T* f) {
    if (vel >= T(0)) return (T(1.) / T(6.)) * f[-2] - f[-1] + (T(1.) / T(2.)) * f[0] + (T(1.) / T(3.)) * f[1];
    return -(T(1.) / T(6.)) * f[2] + f[1] - (T(1.) / T(2.)) * f[0] - (T(1.) / T(3.)) * f[-1];
}

// first-order upwind at k = 1 and k = kmax (hyperPDE.f90:477-481)
template <typename T>
__device__ __forceinline__ T upwind1(T vel, const T* f) {
    if (vel >= T(0)) return f[0] - f[-1];
    return f[1] - f[0];
}

// The quotient a / b.  The FP64 instantiation divides, with the reference's rounding;
// the float one multiplies by the hardware reciprocal of b (v_rcp_f32, 1 ulp), about a
// tenth of the instructions of an IEEE division.  The mixed modes bound their error
// against FP64 instead of matching it bit for bit, and a repeated divisor's
// reciprocal is formed once.
template <typename T>
__device__ __forceinline__ T qdiv(T a, T b) { return a / b; }
template <>
__device__ __forceinline__ float qdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Jiang-Shu WENO5 reconstruction from five differences (hyperPDE.f90:551-576)
template <typename T>
__device__ __forceinline__ T weno5(T v1, T v2, T v3, T v4, T v5) {
    const T eps_weno = T(1.e-6), sigma_1 = T(0.1), sigma_2 = T(0.6), sigma_3 = T(0.3);
    T t1 = v1 - T(2.) * v2 + v3, t2 = v1 - T(4.) * v2 + T(3.) * v3;
    T S1 = (T(13.) / T(12.)) * (t1 * t1) + (T(1.) / T(4.)) * (t2 * t2);
    T t3 = v2 - T(2.) * v3 + v4, t4 = v2 - v4;
    T S2 = (T(13.) / T(12.)) * (t3 * t3) + (T(1.) / T(4.)) * (t4 * t4);
    T t5 = v3 - T(2.) * v4 + v5, t6 = T(3.) * v3 - T(4.) * v4 + v5;
    T S3 = (T(13.) / T(12.)) * (t5 * t5) + (T(1.) / T(4.)) * (t6 * t6);
    T e1 = eps_weno + S1, e2 = eps_weno + S2, e3 = eps_weno + S3;
    T a1 = qdiv(sigma_1, e1 * e1);
    T a2 = qdiv(sigma_2, e2 * e2);
    T a3 = qdiv(sigma_3, e3 * e3);
    T w1 = qdiv(a1, a1 + a2 + a3);
    T w2 = qdiv(a2, a1 + a2 + a3);
    T w3 = qdiv(a3, a1 + a2 + a3);
    return w1 * ((T(1.) / T(3.)) * v1 - (T(7.) / T(6.)) * v2 + (T(11.) / T(6.)) * v3) +
           w2 * ((-T(1.) / T(6.)) * v2 + (T(5.) / T(6.)) * v3 + (T(1.) / T(3.)) * v4) +
           w3 * ((T(1.) / T(3.)) * v3 + (T(5.) / T(6.)) * v4 - (T(1.) / T(6.)) * v5);
}

// Division by a grid spacing.  When the spacing is a power of two (P2) its
// reciprocal is exact and a/h == a*(1/h) bit for bit, so the product is used.
template <typename T, bool P2>
struct Sp {
    T h, hi;
    __device__ __forceinline__ T div(T a) const { return P2 ? a * hi : a / h; }
};

// left-biased derivative (hyperPDE.f90:545-549)
template <typename T, bool P2>
__device__ __forceinline__ T weno_minus(const T* f, Sp<T, P2> s) {
    return weno5<T>(s.div(f[-2] - f[-3]), s.div(f[-1] - f[-2]), s.div(f[0] - f[-1]), s.div(f[1] - f[0]),
                    s.div(f[2] - f[1]));
}
// right-biased derivative (hyperPDE.f90:579-583)
template <typename T, bool P2>
__device__ __forceinline__ T weno_plus(const T* f, Sp<T, P2> s) {
    return weno5<T>(s.div(f[3] - f[2]), s.div(f[2] - f[1]), s.div(f[1] - f[0]), s.div(f[0] - f[-1]),
                    s.div(f[-1] - f[-2]));
}

// seven equal values: every difference is +0, both WENO5 derivatives are exactly +0
template <typename T>
__device__ __forceinline__ bool flat7(const T* f) {
    return f[0] == f[3] && f[1] == f[3] && f[2] == f[3] && f[4] == f[3] && f[5] == f[3] && f[6] == f[3];
}

// Heaviside 'trig' (misc.f90:439-459); pi = acos(-1.)
template <typename T>
__device__ __forceinline__ T heaviside_trig(T phi, T eps, T pi) {
    if (phi < -eps) return T(0);
    if (phi < eps) return T(.5) + T(.5) * phi / eps + T(.5) / pi * sin(pi * phi / eps);
    return T(1);
}

// Dirac 'trig' (misc.f90:518-538)
template <typename T>
__device__ __forceinline__ T dirac_trig(T phi, T eps, T pi) {
    if (phi < -eps) return T(0);
    if (phi < eps) return T(.5) / eps + T(.5) / eps * cos(pi * phi / eps);
    return T(0);
}

// icurvature, 3D branch (interface.f90:126-179); ph/cv point at the centre value
template <typename T>
__device__ __forceinline__ T icurvature(int nb, const T* ph, const T* cv) {
    T sp = T(0);
    if (cv[0] < T(0)) sp = T(-1.);
    if (cv[0] > T(0)) sp = T(1.);
    T eps_0 = sp * ph[0];
    T eps_nb = sp * ph[nb];
    if (cv[0] * cv[nb] > T(0)) {
        T a1 = eps_nb * cv[0];
        T a2 = eps_0 * cv[nb];
        return (a1 - a2) / (eps_nb - eps_0);
    }
    return T(0);
}

// get_curvature's arithmetic (interface.f90:59-109) on 27 gathered values v:
// gradient and Hessian from the least-squares rows of A (10 x 27, column-major),
// kappa = -div(grad phi / |grad phi|), zero when |grad phi| < 0.9, clamped to
// +-max_kappa, zero below 1e-9.  m receives |grad phi|.
template <class T>
__device__ __forceinline__ T curv_lsq(const T* __restrict__ A, const T* v, T max_kappa, T& m) {
    auto row = [&](int r) {
        T s = T(0.);
#pragma unroll
        for (int q = 0; q < 27; q++) s = s + A[r + 10 * q] * v[q];
        return s;
    };
    T dphidx = row(1), dphidy = row(2), dphidz = row(3);
    T d2phidx = row(4), d2phidy = row(5), d2phidz = row(6);
    T d2phidxdy = row(7), d2phidxdz = row(8), d2phidydz = row(9);
    T norm_val = sqrt(dphidx * dphidx + dphidy * dphidy + dphidz * dphidz);
    m = norm_val;
    T cv;
    if (norm_val < T(0.9)) {
        cv = T(0.);
    } else {
        T term1 = (d2phidx + d2phidy + d2phidz) / norm_val;
        T n3 = norm_val * norm_val * norm_val;
        T term2 = dphidx * dphidx * d2phidx + dphidy * dphidy * d2phidy + dphidz * dphidz * d2phidz;
        term2 = term2 / n3;
        T term3 = dphidx * dphidy * d2phidxdy + dphidx * dphidz * d2phidxdz + dphidy * dphidz * d2phidydz;
        term3 = T(2.) * term3 / n3;
        cv = -(term1 - term2 - term3);
    }
    if (cv < -max_kappa) cv = -max_kappa;
    else if (cv > max_kappa) cv = max_kappa;
    else if (fabs(cv) < T(1e-9)) cv = T(0.);
    return cv;
}

}  // namespace icls
