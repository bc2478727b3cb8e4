// qlskum_math.h -- per-point q-LSKUM arithmetic shared by the HIP kernels.
//
// Every routine keeps the reference's expression order (Fortran a+b+c is
// (a+b)+c, a*b*c is (a*b)*c) and the FP64 kernels are compiled with
// -ffp-contract=off, so the only differences from the reference CPU build
// come from the device exp/log/erf (ocml vs glibc, within an ulp or two) --
// and, in the sync-free sweeps, from the sweep form of the split fluxes
// (SWEEP below: two roots from one reciprocal square root, within 2 ulp)
// and from kernels_sweep.hip's -ffp-contract=on (a*b + c fused within an
// expression, identically at every call site).
// Routines are templated on the arithmetic type T so the mixed-precision
// kernels reuse them with float, where division, square root and erf take
// the faster forms defined below.
#pragma once

#include <hip/hip_runtime.h>

namespace mfcfd {

// limiters.f90:26,28 compare |del_neg| with the REAL(4) literal 10e-6
constexpr double kLimiterThreshold = (double)10e-6f;

template <typename T> __device__ __forceinline__ T dexp(T x);
template <> __device__ __forceinline__ double dexp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float dexp<float>(float x) { return __expf(x); }
template <typename T> __device__ __forceinline__ T dlog(T x);
template <> __device__ __forceinline__ double dlog<double>(double x) { return log(x); }
template <> __device__ __forceinline__ float dlog<float>(float x) { return __logf(x); }
template <typename T> __device__ __forceinline__ T dabs(T x) { return x < T(0) ? -x : x; }

// Division, square root and error function of the arithmetic type T.  FP64
// keeps IEEE division and square root and the library erf: the reference's
// arithmetic.  The mixed form's FP32 takes the hardware reciprocal and
// square root (v_rcp_f32, v_sqrt_f32: 1 ulp) and its own erf below; every
// argument here is a normal number.
template <typename T> __device__ __forceinline__ T ddiv(T a, T b);
template <> __device__ __forceinline__ double ddiv<double>(double a, double b) { return a / b; }
template <> __device__ __forceinline__ float ddiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <typename T> __device__ __forceinline__ T dsqrt_fast(T x);
template <> __device__ __forceinline__ double dsqrt_fast<double>(double x) { return __builtin_sqrt(x); }
template <> __device__ __forceinline__ float dsqrt_fast<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

// erf(x), given ex = exp(-x*x), which every flux caller forms anyway.
// FP32: for |x| < 0.5 the Maclaurin series of erf to x^13 (truncation
// < 2e-8); above it Abramowitz & Stegun 7.1.26, erf = 1 - t*(a1 + t*(a2 +
// t*(a3 + t*(a4 + t*a5))))*ex with t = 1/(1 + p|x|) (|error| <= 1.5e-7).
// In FP32 arithmetic the pair stays within 4e-7 of erf over all x; both
// branches are evaluated and one selected, so lanes do not diverge.
template <typename T> __device__ __forceinline__ T derf_e(T x, T ex);
// FP64 erf is kept out of line: inlined, the library routine's working
// registers add to everything live around it and take the fused flux kernel
// from ~150 to ~223 VGPRs (2 instead of 3 resident waves per SIMD)
__device__ __attribute__((noinline)) inline double erf_out_of_line(double x) { return erf(x); }
template <> __device__ __forceinline__ double derf_e<double>(double x, double) { return erf_out_of_line(x); }
// SWEEP selects the sweep form of the split fluxes, for kernels whose flux
// evaluations sit on a dependency chain.  In FP64 it
//   * inlines the library erf (registers are spare there: 3% shorter), and
//   * takes sqrt(beta) and 0.5/sqrt(pi*beta) from one reciprocal square root
//     of beta (beta_roots), so that the coefficient B is one multiply after
//     exp instead of an IEEE division (within 2 ulp of the IEEE forms).
// FP32 is the same in both forms.
template <typename T, bool SWEEP>
__device__ __forceinline__ T derf_sel(T x, T ex) {
    if constexpr (SWEEP && sizeof(T) == 8) return erf(x);
    else return derf_e(x, ex);
}

// 1/sqrt(x) to double precision: the hardware estimate (v_rsq_f64) refined
// by two Newton-Raphson steps y <- y*(1.5 - 0.5*x*y*y); x is a positive
// normal number
__device__ __forceinline__ double rsqrt_refined(double x) {
    const double hx = 0.5 * x;
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-hx, y * y, 1.5);
    y = y * fma(-hx, y * y, 1.5);
    return y;
}

// sb = sqrt(beta) and hb = 0.5/sqrt(pi*beta), so that B = exp(-S*S)*hb
template <typename T, bool SWEEP>
__device__ __forceinline__ void beta_roots(T beta, T pi, T& sb, T& hb) {
    if constexpr (SWEEP && sizeof(T) == 8) {
        const T y = rsqrt_refined(beta);
        sb = beta * y;
        hb = (T(0.5) * rsqrt_refined(pi)) * y;
    } else {
        sb = dsqrt_fast(beta);
        hb = T(0);   // unused: the reference form divides by sqrt(pi*beta)
    }
}

template <typename T, bool SWEEP>
__device__ __forceinline__ T b_coef(T ex, T hb, T pi, T beta) {
    if constexpr (SWEEP && sizeof(T) == 8) return ex * hb;
    else return ddiv(T(0.5) * ex, dsqrt_fast(pi * beta));
}
template <> __device__ __forceinline__ float derf_e<float>(float x, float ex) {
    const float ax = x < 0.0f ? -x : x;
    const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * ax);
    const float p =
        t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
    const float far = 1.0f - p * ex;
    // 2/sqrt(pi) * sum_n (-1)^n x^(2n+1) / (n! (2n+1)), n = 0..6
    const float x2 = ax * ax;
    float s = 1.2055332981789664e-4f;
    s = s * x2 - 8.548327023450852e-4f;
    s = s * x2 + 5.223977625442188e-3f;
    s = s * x2 - 2.6866170645131252e-2f;
    s = s * x2 + 1.1283791670955126e-1f;
    s = s * x2 - 3.7612638903183754e-1f;
    s = s * x2 + 1.1283791670955126f;
    const float r = ax < 0.5f ? ax * s : far;
    return x < 0.0f ? -r : r;
}

// The six symmetric sums of a weighted least-squares stencil and the 3x3
// Cramer solve every stencil of the reference uses
// (q_derivatives.f90:85-107, interior_flux_dGxpos.f90:115-123).
template <typename T>
struct Geo6 {
    T xx, yy, zz, xy, yz, zx;
};

template <typename T>
__device__ __forceinline__ T cramer_det(const Geo6<T>& g) {
    return g.xx * (g.yy * g.zz - g.yz * g.yz)
         - g.xy * (g.xy * g.zz - g.yz * g.zx)
         + g.zx * (g.xy * g.yz - g.yy * g.zx);
}

// numerator of solution component col for the right-hand side (a, b, c)
template <typename T>
__device__ __forceinline__ T cramer_num(const Geo6<T>& g, int col, T a, T b, T c) {
    if (col == 0)
        return a * (g.yy * g.zz - g.yz * g.yz)
             - b * (g.xy * g.zz - g.zx * g.yz)
             + c * (g.xy * g.yz - g.zx * g.yy);
    if (col == 1)
        return g.xx * (b * g.zz - g.yz * c)
             - g.xy * (a * g.zz - g.zx * c)
             + g.zx * (a * g.yz - g.zx * b);
    return g.xx * (g.yy * c - b * g.yz)
         - g.xy * (g.xy * c - a * g.yz)
         + g.zx * (g.xy * b - a * g.yy);
}

template <typename T>
struct Frame {
    T t1[3], t2[3], n[3];
};

// Half-range KFVS flux of split_fluxes.f90:11-308 along axis 0/1/2
// (tan1/tan2/nor), half-space sign +1/-1.  "x - y" of the negative variants
// is written "x + (-y)", which rounds identically.
template <typename T, bool SWEEP = false>
__device__ __forceinline__ void flux_half(int axis, int sign, const Frame<T>& f, const T prim[5],
                                          T pi, T G[5]) {
    const T rho = prim[0], u1 = prim[1], u2 = prim[2], u3 = prim[3], pr = prim[4];
    const T ut1 = f.t1[0] * u1 + f.t1[1] * u2 + f.t1[2] * u3;
    const T ut2 = f.t2[0] * u1 + f.t2[1] * u2 + f.t2[2] * u3;
    const T un = f.n[0] * u1 + f.n[1] * u2 + f.n[2] * u3;
    const T beta = ddiv(T(0.5) * rho, pr);
    const T ua = axis == 0 ? ut1 : (axis == 1 ? ut2 : un);
    T sb, hb;
    beta_roots<T, SWEEP>(beta, pi, sb, hb);
    const T S = ua * sb;
    const T ex = dexp(-S * S);
    const T B = b_coef<T, SWEEP>(ex, hb, pi, beta);
    const T e = derf_sel<T, SWEEP>(S, ex);
    const T A = T(0.5) * (sign > 0 ? T(1) + e : T(1) - e);
    const T Bs = sign > 0 ? B : -B;
    const T u_sqr = ut1 * ut1 + ut2 * ut2 + un * un;
    T temp1 = (ua * A + Bs);
    G[0] = rho * temp1;
    G[1] = rho * ut1 * temp1;
    G[2] = rho * ut2 * temp1;
    G[3] = rho * un * temp1;
    G[1 + axis] = pr * A + rho * ua * temp1;
    temp1 = T(2.5) * pr + T(0.5) * rho * u_sqr;
    const T tail = (temp1 + T(0.5) * pr) * B;
    G[4] = (temp1 + pr) * ua * A + (sign > 0 ? tail : -tail);
}

// Quarter-range fluxes of octant_fluxes.f90:9-442: axis 0/1, sign_a along
// that axis, sign_n = -1 wall family (A3neg, -B3), +1 outer family (A3pos).
template <typename T, bool SWEEP = false>
__device__ __forceinline__ void flux_octant(int axis, int sign_a, int sign_n, const Frame<T>& f,
                                            const T prim[5], T pi, T G[5]) {
    const T rho = prim[0], u1 = prim[1], u2 = prim[2], u3 = prim[3], pr = prim[4];
    const T ut1 = f.t1[0] * u1 + f.t1[1] * u2 + f.t1[2] * u3;
    const T ut2 = f.t2[0] * u1 + f.t2[1] * u2 + f.t2[2] * u3;
    const T un = f.n[0] * u1 + f.n[1] * u2 + f.n[2] * u3;
    const T beta = ddiv(T(0.5) * rho, pr);
    const T ua = axis == 0 ? ut1 : ut2;
    T sb, hb;
    beta_roots<T, SWEEP>(beta, pi, sb, hb);
    const T Sa = ua * sb;
    const T S3 = un * sb;
    const T exa = dexp(-Sa * Sa), ex3 = dexp(-S3 * S3);
    const T Ba = b_coef<T, SWEEP>(exa, hb, pi, beta);
    const T B3 = b_coef<T, SWEEP>(ex3, hb, pi, beta);
    const T ea = derf_sel<T, SWEEP>(Sa, exa);
    const T Aa = T(0.5) * (sign_a > 0 ? T(1) + ea : T(1) - ea);
    const T e3 = derf_sel<T, SWEEP>(S3, ex3);
    const T A3 = T(0.5) * (sign_n > 0 ? T(1) + e3 : T(1) - e3);
    const T u_sqr = ut1 * ut1 + ut2 * ut2 + un * un;
    const T temp1 = (ua * Aa + (sign_a > 0 ? Ba : -Ba));
    G[0] = rho * temp1 * A3;
    if (axis == 0) {
        G[1] = (pr * Aa + rho * ut1 * temp1) * A3;
        G[2] = rho * ut2 * temp1 * A3;
    } else {
        G[1] = rho * ut1 * temp1 * A3;
        G[2] = (pr * Aa + rho * ut2 * temp1) * A3;
    }
    G[3] = rho * temp1 * (un * A3 + (sign_n > 0 ? B3 : -B3));
    const T temp2 = T(2.5) * pr + T(0.5) * rho * u_sqr;
    const T tb = (temp2 + T(0.5) * pr) * Ba;
    const T temp3 = (temp2 + pr) * ua * Aa + (sign_a > 0 ? tb : -tb);
    const T tn = T(0.5) * rho * un * B3 * temp1;
    G[4] = temp3 * A3 + (sign_n > 0 ? tn : -tn);
}

// flux family selector: 0 half-range, 1 wall octant (Gw), 2 outer octant (Go)
template <typename T, bool SWEEP = false>
__device__ __forceinline__ void flux_eval(int family, int axis, int sign, const Frame<T>& f,
                                          const T prim[5], T pi, T G[5]) {
    if (family == 0)
        flux_half<T, SWEEP>(axis, sign, f, prim, pi, G);
    else
        flux_octant<T, SWEEP>(axis, sign, family == 2 ? 1 : -1, f, prim, pi, G);
}

// limiters.f90:18-40 for one component, from del_neg = qtilde - q
template <typename T>
__device__ __forceinline__ T venkat_phi_dn(T del_neg, T q, T qmax, T qmin, T epsi2) {
    const T thr = T(kLimiterThreshold);
    if (dabs(del_neg) <= thr) return T(1);
    if (!(dabs(del_neg) > thr)) return T(1);  // NaN: the reference leaves phi unset
    const T del_pos = del_neg > T(0) ? qmax - q : qmin - q;
    T num = (del_pos * del_pos) + epsi2;
    num = num * del_neg + T(2) * del_neg * del_neg * del_pos;
    T den = del_pos * del_pos + T(2) * del_neg * del_neg;
    den = den + del_neg * del_pos + epsi2;
    den = den * del_neg;
    const T temp = ddiv(num, den);
    return temp < T(1) ? temp : T(1);
}

template <typename T>
__device__ __forceinline__ T venkat_phi(T qt, T q, T qmax, T qmin, T epsi2) {
    return venkat_phi_dn(qt - q, q, qmax, qmin, epsi2);
}

// qvariables_to_primitive_variables.f90:16-30
template <typename T>
__device__ __forceinline__ void qtilde_to_primitive(const T q[5], T prim[5]) {
    const T beta = -q[4] * T(0.5);
    const T temp = ddiv(T(0.5), beta);
    prim[1] = q[1] * temp;
    prim[2] = q[2] * temp;
    prim[3] = q[3] * temp;
    const T temp1 = q[0] + beta * (prim[1] * prim[1] + prim[2] * prim[2] + prim[3] * prim[3]);
    const T temp2 = temp1 - (dlog(beta) * T(2.5));
    prim[0] = dexp(temp2);
    prim[4] = prim[0] * temp;
}

// get_primitive_vector.f90:10-41 / conserved_to_primitive_vector.f90:10-42:
// local-frame conserved vector of a point -> global primitive vector
template <typename T>
__device__ __forceinline__ void conserved_to_prim(const T U[5], const Frame<T>& f, T prim[5]) {
    prim[0] = U[0];
    T temp = T(1) / U[0];
    const T u2 = f.t1[0] * U[1] + f.t2[0] * U[2] + f.n[0] * U[3];
    const T u3 = f.t1[1] * U[1] + f.t2[1] * U[2] + f.n[1] * U[3];
    const T u4 = f.t1[2] * U[1] + f.t2[2] * U[2] + f.n[2] * U[3];
    prim[1] = temp * u2;
    prim[2] = temp * u3;
    prim[3] = temp * u4;
    temp = prim[1] * prim[1] + prim[2] * prim[2] + prim[3] * prim[3];
    prim[4] = T(0.4) * (U[4] - T(0.5) * prim[0] * temp);
}

}  // namespace mfcfd
