// physics.hpp -- per-face and per-zone physics of Mara2's Godunov step as
// device functions, templated on the working precision R (double for the
// plain path, float for the flux path of the mixed form).
//
// Every routine keeps the operation order of the reference routine it cites,
// so that with contraction disabled (-ffp-contract=off) the FP64 results are
// the reference's bit for bit.  NaN behaviour of the reference's comparison
// macros is kept as well (they are written as ternaries, never fmin/fmax).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <type_traits>

namespace mara {

// ---------------------------------------------------------------------------
// precision helpers
// ---------------------------------------------------------------------------
template <typename R> struct Math;

template <> struct Math<double>
{
    static __device__ __forceinline__ double sqrt(double x) { return __builtin_sqrt(x); }
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
    // a / b where b is positive and finite (the same division here)
    static __device__ __forceinline__ double div_pos(double a, double b) { return a / b; }
    static __device__ __forceinline__ double abs(double x) { return __builtin_fabs(x); }
    // the reference's MIN3 / MAX3 and comparison order (NaN semantics kept)
    static __device__ __forceinline__ double min3(double a, double b, double c)
    {
        return a < b ? (a < c ? a : c) : (b < c ? b : c);
    }
    static __device__ __forceinline__ double max3(double a, double b, double c)
    {
        return a > b ? (a > c ? a : c) : (b > c ? b : c);
    }
    static constexpr bool ieee_minmax = false;
};

// binary32 in the mixed form: hardware square root and reciprocal (about one
// ulp), which the mixed form's L1 budget of 1e-5 against FP64 absorbs.
template <> struct Math<float>
{
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float div_pos(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float abs(float x) { return __builtin_fabsf(x); }
    // minima and maxima as single min/max instructions (the comparison
    // forms differ only where an operand is NaN, a state already lost)
    static __device__ __forceinline__ float min3(float a, float b, float c)
    {
        return __builtin_fminf(a, __builtin_fminf(b, c));
    }
    static __device__ __forceinline__ float max3(float a, float b, float c)
    {
        return __builtin_fmaxf(a, __builtin_fmaxf(b, c));
    }
    static constexpr bool ieee_minmax = true;
};

// Binary64 arithmetic of the relativistic law.  Its results are held to
// relative L1 1e-12 of the reference, not to the reference's rounding (its
// quartic already goes through the device libm), so its divisions and
// square roots take the hardware reciprocal (reciprocal square root, relative
// error up to ~5e-8) refined by one Newton step, within 2.1e-15 (4.1e-15),
// measured on gfx950 over 2^20 log-uniform inputs, in place of the correctly
// rounded sequences (about ten instructions each), and its expressions
// contract to fused multiply-adds where they are written as multiply-adds.
struct Rel
{
    // 1/x; x = 0, +-inf and NaN give the hardware value (+-inf, 0, NaN)
    static __device__ __forceinline__ double rcp(double x)
    {
        const double y0 = __builtin_amdgcn_rcp(x);
        const double y = __builtin_fma(y0, __builtin_fma(-x, y0, 1.0), y0);
        return __builtin_isfinite(y) ? y : y0;
    }
    static __device__ __forceinline__ double div(double a, double b) { return a * rcp(b); }
    // a / b for b known positive and finite (WENO5's weights, HLLE's wave-speed
    // spread): the refined reciprocal without the zero and infinity fix-up,
    // which such b never needs (a zero spread gives NaN either way)
    static __device__ __forceinline__ double div_pos(double a, double b)
    {
        const double y0 = __builtin_amdgcn_rcp(b);
        return a * __builtin_fma(y0, __builtin_fma(-b, y0, 1.0), y0);
    }
    // sqrt(x) as x / sqrt(x); 0 and +inf as they are, NaN below 0
    static __device__ __forceinline__ double sqrt(double x)
    {
        const double r0 = __builtin_amdgcn_rsq(x);
        const double r = __builtin_fma(0.5 * r0, __builtin_fma(-(x * r0), r0, 1.0), r0);
        return (x == 0.0 || x == __builtin_inf()) ? x : x * r;
    }
    static __device__ __forceinline__ double abs(double x) { return __builtin_fabs(x); }
    // minima and maxima as single min/max instructions; they differ from the
    // comparison forms only where an operand is NaN (or for zeros of both
    // signs, which these operands never are), a state already lost
    static __device__ __forceinline__ double min3(double a, double b, double c)
    {
        return __builtin_fmin(a, __builtin_fmin(b, c));
    }
    static __device__ __forceinline__ double max3(double a, double b, double c)
    {
        return __builtin_fmax(a, __builtin_fmax(b, c));
    }
    static constexpr bool ieee_minmax = true;
};

// the relativistic law's arithmetic in working precision R: Rel in
// binary64, the mixed form's binary32 one otherwise
template <typename R> struct RelMath { using type = Math<R>; };
template <> struct RelMath<double> { using type = Rel; };

// A law's arithmetic by default: Rel for the relativistic law (its binary64
// form is held to a bound, not to the reference's rounding), the working
// precision's own for the others.  The fast binary64 form passes Rel for the
// MHD law as well.
template <int Law, typename R>
using LawMath = typename std::conditional<Law == 2, typename RelMath<R>::type, Math<R>>::type;

// max / min of two values as one instruction in the working precision
template <typename R> __device__ __forceinline__ R fmax2(R a, R b)
{
    if constexpr (sizeof(R) == 4) return __builtin_fmaxf(a, b);
    else return __builtin_fmax(a, b);
}
template <typename R> __device__ __forceinline__ R fmin2(R a, R b)
{
    if constexpr (sizeof(R) == 4) return __builtin_fminf(a, b);
    else return __builtin_fmin(a, b);
}

template <typename R> __device__ __forceinline__ R min2(R a, R b) { return a < b ? a : b; }
template <typename R> __device__ __forceinline__ R max2(R a, R b) { return a > b ? a : b; }
// MIN3 / MAX3 of src/RiemannSolvers.cpp:3-6
template <typename R> __device__ __forceinline__ R min3(R a, R b, R c) { return a < b ? min2(a, c) : min2(b, c); }
template <typename R> __device__ __forceinline__ R max3(R a, R b, R c) { return a > b ? max2(a, c) : max2(b, c); }
template <typename R> __device__ __forceinline__ int sgn(R x) { return (x > R(0)) - (x < R(0)); }

// ---------------------------------------------------------------------------
// Reconstruction (src/Reconstruction.cpp:55-132)
// ---------------------------------------------------------------------------
template <typename R, typename A = Math<R>>
__device__ __forceinline__ R minmod(R theta, R ul, R u0, R ur)
{
    const R a = theta * (u0 - ul);
    const R b = R(0.5) * (ur - ul);
    const R c = theta * (ur - u0);
    if constexpr (sizeof(R) == 4 || A::ieee_minmax)
    {
        // the mixed and fast forms: the same value (the argument of least
        // magnitude when all three share a sign, else zero) as min/max
        // instructions, max(min(a, b, c), min(max(a, b, c), 0)); only the
        // sign of a zero result can differ from the reference's product form
        const R lo = fmin2(a, fmin2(b, c));
        const R hi = fmax2(a, fmax2(b, c));
        return fmax2(lo, fmin2(hi, R(0)));
    }
    const R m = min3(Math<R>::abs(a), Math<R>::abs(b), Math<R>::abs(c));
    const int sa = sgn(a), sb = sgn(b), sc = sgn(c);
    return R(0.25) * Math<R>::abs(R(sa + sb)) * R(sa + sc) * m;
}

// PLM_C2R (sgn=+1) / PLM_C2L (sgn=-1): v0 + sgn * 0.5 * minmod
template <typename R, typename A = Math<R>>
__device__ __forceinline__ R plm(R theta, R vl, R v0, R vr, R sgnv)
{
    return v0 + sgnv * R(0.5) * minmod<R, A>(theta, vl, v0, vr);
}

// weno5 with the FD tables and the ImprovedShenZha10 indicator, A = 50
// (src/Reconstruction.cpp:68-132, 154-169).  Right = true selects C2R
// (value at the right edge of the centre cell), else C2L.  v[-2..2].  D
// supplies the division (Math<R>, or the relativistic law's Rel).
template <typename R, bool Right, typename D = Math<R>>
__device__ __forceinline__ R weno5(R vm2, R vm1, R v0, R vp1, R vp2)
{
    // coefficient rows c[r][0..2] and weights d[r]
    const R c00 = Right ? R(1. / 3.) : R(11. / 6.), c01 = Right ? R(5. / 6.) : R(-7. / 6.), c02 = Right ? R(-1. / 6.) : R(1. / 3.);
    const R c10 = Right ? R(-1. / 6.) : R(1. / 3.), c11 = R(5. / 6.), c12 = Right ? R(1. / 3.) : R(-1. / 6.);
    const R c20 = Right ? R(1. / 3.) : R(-1. / 6.), c21 = Right ? R(-7. / 6.) : R(5. / 6.), c22 = Right ? R(11. / 6.) : R(1. / 3.);
    const R d0 = Right ? R(0.3) : R(0.1), d1 = R(0.6), d2 = Right ? R(0.1) : R(0.3);
    const R eps_prime = R(1e-10);
    const R A = R(50.0);
    const R vs0 = c00 * v0 + c01 * vp1 + c02 * vp2;
    const R vs1 = c10 * vm1 + c11 * v0 + c12 * vp1;
    const R vs2 = c20 * vm2 + c21 * vm1 + c22 * v0;
    const R t0a = R(1) * v0 - R(2) * vp1 + R(1) * vp2;
    const R t0b = R(3) * v0 - R(4) * vp1 + R(1) * vp2;
    const R t1a = R(1) * vm1 - R(2) * v0 + R(1) * vp1;
    const R t1b = R(1) * vm1 - R(0) * v0 - R(1) * vp1;
    const R t2a = R(1) * vm2 - R(2) * vm1 + R(1) * v0;
    const R t2b = R(1) * vm2 - R(4) * vm1 + R(3) * v0;
    R B0 = R(13. / 12.) * (t0a * t0a) + R(1. / 4.) * (t0b * t0b);
    R B1 = R(13. / 12.) * (t1a * t1a) + R(1. / 4.) * (t1b * t1b);
    R B2 = R(13. / 12.) * (t2a * t2a) + R(1. / 4.) * (t2b * t2b);
    const R minB = D::min3(B0, B1, B2);
    const R maxB = D::max3(B0, B1, B2);
    const R R0 = D::div_pos(minB, maxB + eps_prime);
    B0 = R0 * A * minB + B0;
    B1 = R0 * A * minB + B1;
    B2 = R0 * A * minB + B2;
    const R e0 = eps_prime + B0, e1 = eps_prime + B1, e2 = eps_prime + B2;
    const R w0 = D::div_pos(d0, e0 * e0);
    const R w1 = D::div_pos(d1, e1 * e1);
    const R w2 = D::div_pos(d2, e2 * e2);
    const R wtot = w0 + w1 + w2;
    return D::div_pos(w0 * vs0 + w1 * vs1 + w2 * vs2, wtot);
}

// ---------------------------------------------------------------------------
// Real roots of a quartic (the job of QuarticPolynomial::solve,
// src/QuarticPolynomial.cpp:19-144), binary64 only, by Ferrari's method.
// ---------------------------------------------------------------------------
// The largest real root of y^3 + A y^2 + B y + C = 0.  With y = t - A/3 the
// cubic is t^3 + 3 h t - 2 g = 0 and disc = h^3 + g^2 decides the case:
// one real root (disc > 0, Cardano: t = cbrt(g + sqrt(disc)) + cbrt(g -
// sqrt(disc))), three (disc < 0, Viete: the largest is 2 sqrt(-h)
// cos(phi / 3) with cos(phi) = g / sqrt(-h)^3), or a repeated one (disc = 0:
// t = 2 cbrt(g) and the double root -cbrt(g)).
__device__ inline double cubic_largest_root(double A, double B, double C)
{
#pragma clang fp contract(fast)
    const double shift = Rel::div_pos(A, 3.0);
    const double h = Rel::div_pos(B, 3.0) - shift * shift;
    const double g = 0.5 * (Rel::div_pos(A * B, 3.0) - C) - shift * shift * shift;
    const double disc = h * h * h + g * g;
    double t;
    if (disc > 0.0)
    {
        const double w = Rel::sqrt(disc);
        t = cbrt(g + w) + cbrt(g - w);
    }
    else if (disc < 0.0)
    {
        const double m = Rel::sqrt(-h);
        double c = Rel::div_pos(g, m * m * m);
        c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
        t = 2.0 * m * cos(Rel::div_pos(acos(c), 3.0));
    }
    else
    {
        const double c = cbrt(g);
        t = 2.0 * c > -c ? 2.0 * c : -c;
    }
    return t - shift;
}

// Real roots of k4 x^4 + k3 x^3 + k2 x^2 + k1 x + k0 = 0, ascending in
// roots[]; returns their number, 0, 2 or 4.  For the monic x^4 + b x^3 +
// c x^2 + d x + e take the largest root y of the resolvent cubic
// y^3 - c y^2 + (b d - 4 e) y + (4 c e - d^2 - b^2 e) = 0; then
// x^4 + ... = (x^2 + b x / 2 + y / 2)^2 - (s x + u)^2 with
// s^2 = b^2 / 4 - c + y, and the quartic is the product of the quadratics
// x^2 + (b/2 - s) x + ... and x^2 + (b/2 + s) x + ..., whose roots are
// -b/4 + s/2 +- sqrt(q+)/2 and -b/4 - s/2 +- sqrt(q-)/2 with
// q+- = 3 b^2 / 4 - 2 c - s^2 +- (b c - 2 d - b^3 / 4) / s.  Where s = 0 the
// linear term vanishes and q+- = 3 b^2 / 4 - 2 c +- 2 sqrt(y^2 - 4 e).  A
// quadratic with a negative q contributes no real root.
__device__ inline int quartic_real_roots(double k4, double k3, double k2, double k1, double k0, double roots[4])
{
#pragma clang fp contract(fast)
    const double ik4 = Rel::rcp(k4);
    const double b = k3 * ik4, c = k2 * ik4, d = k1 * ik4, e = k0 * ik4;
    const double y = cubic_largest_root(-c, b * d - 4.0 * e, 4.0 * c * e - d * d - b * b * e);
    const double s2 = 0.25 * b * b + y - c;
    const double s = s2 > 0.0 ? Rel::sqrt(s2) : 0.0;
    const double common = 0.75 * b * b - 2.0 * c;
    double qp, qm;
    if (s != 0.0)
    {
        const double lin = Rel::div_pos(0.25 * (4.0 * b * c - 8.0 * d - b * b * b), s);
        qp = common - s2 + lin;
        qm = common - s2 - lin;
    }
    else
    {
        const double w = 2.0 * Rel::sqrt(y * y - 4.0 * e);
        qp = common + w;
        qm = common - w;
    }
    const double centre = -0.25 * b;
    int n = 0;
    if (qp >= 0.0)
    {
        const double half = 0.5 * Rel::sqrt(qp);
        roots[n++] = centre + 0.5 * s - half;
        roots[n++] = centre + 0.5 * s + half;
    }
    if (qm >= 0.0)
    {
        const double half = 0.5 * Rel::sqrt(qm);
        roots[n++] = centre - 0.5 * s - half;
        roots[n++] = centre - 0.5 * s + half;
    }
    if (n == 4)
    {
        // two ascending pairs merged by compare-exchange: (0,2), (1,3), (1,2)
        auto order = [&](int i, int j) {
            if (roots[j] < roots[i])
            {
                const double x = roots[i];
                roots[i] = roots[j];
                roots[j] = x;
            }
        };
        order(0, 2);
        order(1, 3);
        order(1, 2);
    }
    return n;
}

// ---------------------------------------------------------------------------
// Face states.  A State holds what the Riemann solvers read: P, U, F and the
// eigenvalues A (unused slots of the reference's 10-wide arrays are zero).
// ---------------------------------------------------------------------------
template <typename R, int NQ>
struct State
{
    R P[NQ], U[NQ], F[NQ], A[NQ];
};

enum : int { ERR_NONE = 0, ERR_SUPERLUMINAL = 1, ERR_EIGEN = 2 };

// NewtonianHydro::fromPrimitive (src/ConservationLaws.cpp:253-290)
template <typename R>
__device__ __forceinline__ void hydro_state(R gamma, const R* P, const R* dAA, State<R, 5>& S)
{
    const R gm0 = gamma;
    const R gm1 = gamma - R(1.0);
    const R cs = Math<R>::sqrt(Math<R>::div(gm0 * P[4], P[0]));
    const R vv = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
    const R vn = P[1] * dAA[0] + P[2] * dAA[1] + P[3] * dAA[2];
#pragma unroll
    for (int q = 0; q < 5; ++q) S.P[q] = P[q];
    S.U[0] = P[0];
    S.U[1] = P[0] * P[1];
    S.U[2] = P[0] * P[2];
    S.U[3] = P[0] * P[3];
    S.U[4] = P[0] * R(0.5) * vv + Math<R>::div(P[4], gm1);
    S.F[0] = vn * S.U[0];
    S.F[1] = vn * S.U[1] + P[4] * dAA[0];
    S.F[2] = vn * S.U[2] + P[4] * dAA[1];
    S.F[3] = vn * S.U[3] + P[4] * dAA[2];
    S.F[4] = vn * S.U[4] + P[4] * vn;
    S.A[0] = vn - cs;
    S.A[1] = vn;
    S.A[2] = vn;
    S.A[3] = vn;
    S.A[4] = vn + cs;
}

// NewtonianMHD::fromPrimitive (src/ConservationLaws.cpp:416-487)
// A: the arithmetic (Math<R>, or Rel in the fast binary64 form)
template <typename R, typename A = Math<R>>
__device__ __forceinline__ void mhd_state(R gamma, const R* P, const R* dAA, State<R, 8>& S)
{
    const R gm0 = gamma;
    const R gm1 = gamma - R(1.0);
    const R cs2 = A::div_pos(gm0 * P[4], P[0]);
    const R vv = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
    const R BB = P[5] * P[5] + P[6] * P[6] + P[7] * P[7];
    const R Bv = P[5] * P[1] + P[6] * P[2] + P[7] * P[3];
    const R vn = P[1] * dAA[0] + P[2] * dAA[1] + P[3] * dAA[2];
    const R Bn = P[5] * dAA[0] + P[6] * dAA[1] + P[7] * dAA[2];
    const R ps = P[4] + R(0.5) * BB;
#pragma unroll
    for (int q = 0; q < 8; ++q) S.P[q] = P[q];
    S.U[0] = P[0];
    S.U[1] = P[0] * P[1];
    S.U[2] = P[0] * P[2];
    S.U[3] = P[0] * P[3];
    S.U[4] = P[0] * R(0.5) * vv + R(0.5) * BB + A::div_pos(P[4], gm1);
    S.U[5] = P[5];
    S.U[6] = P[6];
    S.U[7] = P[7];
    S.F[0] = vn * S.U[0];
    S.F[1] = vn * S.U[1] - Bn * P[5] + ps * dAA[0];
    S.F[2] = vn * S.U[2] - Bn * P[6] + ps * dAA[1];
    S.F[3] = vn * S.U[3] - Bn * P[7] + ps * dAA[2];
    S.F[4] = vn * S.U[4] - Bn * Bv + ps * vn;
    S.F[5] = vn * S.U[5] - Bn * P[1];
    S.F[6] = vn * S.U[6] - Bn * P[2];
    S.F[7] = vn * S.U[7] - Bn * P[3];
    const R Bn2 = Bn * Bn;
    const R ca2 = A::div_pos(BB, P[0]);
    const R cn2 = A::div_pos(Bn2, P[0]);
    const R cw4 = (cs2 + ca2) * (cs2 + ca2);
    // (cs2 + ca2)^2 - 4 cs2 cn2 >= (cs2 - ca2)^2 >= 0 since cn2 <= ca2; in
    // binary32, or with contraction, it can round below zero where cs2 ~ ca2
    // and B is along the normal, and the square root would then poison the
    // fast speed.  The mixed and fast forms clamp it; the binary64 form
    // keeps the reference's expression.
    R disc = cw4 - R(4) * cs2 * cn2;
    if constexpr (sizeof(R) == 4 || A::ieee_minmax) disc = disc < R(0) ? R(0) : disc;
    const R root = A::sqrt(disc);
    const R cF2 = R(0.5) * (cs2 + ca2 + root);
    const R cS2 = R(0.5) * (cs2 + ca2 - root);
    const R sF = A::sqrt(cF2), sA = A::sqrt(cn2), sS = A::sqrt(cS2);
    S.A[0] = vn - sF;
    S.A[1] = vn - sA;
    S.A[2] = vn - sS;
    S.A[3] = vn;
    S.A[4] = vn;
    S.A[5] = vn + sS;
    S.A[6] = vn + sA;
    S.A[7] = vn + sF;
}

// RelativisticMHD eigenvalues (src/ConservationLaws.cpp:739-835), binary64.
// Returns ERR_NONE or ERR_EIGEN.
__device__ inline int rmhd_eigen(double gamma, const double* P, const double* dAA, double* A)
{
#pragma clang fp contract(fast)
    const double gm0 = gamma;
    const double gm1 = gamma - 1.0;
    const double vv = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
    const double BB = P[5] * P[5] + P[6] * P[6] + P[7] * P[7];
    const double Bv = P[5] * P[1] + P[6] * P[2] + P[7] * P[3];
    const double vn = P[1] * dAA[0] + P[2] * dAA[1] + P[3] * dAA[2];
    const double W2 = Rel::div_pos(1.0, 1.0 - vv);
    const double W0 = Rel::sqrt(W2);
    const double b0 = W0 * Bv;
    const double bb = Rel::div_pos(BB + b0 * b0, W2);
    const double b1 = Rel::div_pos(P[5] + b0 * W0 * P[1], W0);
    const double b2 = Rel::div_pos(P[6] + b0 * W0 * P[2], W0);
    const double b3 = Rel::div_pos(P[7] + b0 * W0 * P[3], W0);
    const double bn = b1 * dAA[0] + b2 * dAA[1] + b3 * dAA[2];
    const double e0 = Rel::div_pos(Rel::div_pos(P[4], P[0]), gm1);
    const double p0 = P[4];
    const double h0 = 1.0 + e0 + Rel::div_pos(p0, P[0]);
    const double cs2 = Rel::div_pos(gm0 * p0, P[0] * h0);
    const double W4 = W2 * W2;
    const double v2 = vn * vn;
    const double v3 = vn * v2;
    const double v4 = vn * v3;
    if (BB == 0.0)
    {
        const double K = Rel::sqrt(cs2 * (1 - vv) * (1 - vv * cs2 - v2 * (1 - cs2)));
        A[0] = Rel::div_pos(vn * (1 - cs2) - K, 1 - vv * cs2);
        for (int n = 1; n < 7; ++n) A[n] = vn;
        A[7] = Rel::div_pos(vn * (1 - cs2) + K, 1 - vv * cs2);
        return ERR_NONE;
    }
    const double K = W4 * (P[0] * h0 * (Rel::div_pos(1.0, cs2) - 1.));
    const double L = -W2 * (P[0] * h0 + Rel::div_pos(bb, cs2));
    const double A4 = K - L - b0 * b0;
    const double A3 = -4 * K * vn + L * vn * 2 + 2 * b0 * bn;
    const double A2 = 6 * K * v2 + L * (1. - v2) + b0 * b0 - bn * bn;
    const double A1 = -4 * K * v3 - L * vn * 2 - 2 * b0 * bn;
    const double A0 = K * v4 + L * v2 + bn * bn;
    double roots[4];
    const int nr = quartic_real_roots(A4, A3, A2, A1, A0, roots);
    const double C = Rel::sqrt(h0 + bb);
    if (nr == 4)
    {
        A[0] = roots[0];
        A[1] = Rel::div(bn - C * vn * W0, b0 - C * W0);
        A[2] = roots[1];
        A[3] = vn;
        A[4] = vn;
        A[5] = roots[2];
        A[6] = Rel::div(bn + C * vn * W0, b0 + C * W0);
        A[7] = roots[3];
        return ERR_NONE;
    }
    if (nr == 2)
    {
        A[0] = roots[0];
        A[1] = Rel::div(bn - C * vn * W0, b0 - C * W0);
        A[2] = vn;
        A[3] = vn;
        A[4] = vn;
        A[5] = vn;
        A[6] = Rel::div(bn + C * vn * W0, b0 + C * W0);
        A[7] = roots[1];
        return ERR_NONE;
    }
    return ERR_EIGEN;
}

// RelativisticMHD::fromPrimitive, conserved and flux part
// (src/ConservationLaws.cpp:659-732); eigenvalues always in binary64.
template <typename R>
__device__ __forceinline__ int rmhd_state(R gamma, const R* P, const R* dAA, State<R, 8>& S)
{
#pragma clang fp contract(fast)
    using RM = typename RelMath<R>::type;
    const R gm1 = gamma - R(1.0);
    const R vv = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
    const R BB = P[5] * P[5] + P[6] * P[6] + P[7] * P[7];
    const R Bv = P[5] * P[1] + P[6] * P[2] + P[7] * P[3];
    const R vn = P[1] * dAA[0] + P[2] * dAA[1] + P[3] * dAA[2];
    const R Bn = P[5] * dAA[0] + P[6] * dAA[1] + P[7] * dAA[2];
    const R W2 = RM::div_pos(R(1.0), R(1.0) - vv);
    const R W0 = RM::sqrt(W2);
    const R b0 = W0 * Bv;
    const R bb = RM::div_pos(BB + b0 * b0, W2);
    const R b1 = RM::div_pos(P[5] + b0 * W0 * P[1], W0);
    const R b2 = RM::div_pos(P[6] + b0 * W0 * P[2], W0);
    const R b3 = RM::div_pos(P[7] + b0 * W0 * P[3], W0);
    const R e0 = RM::div_pos(RM::div_pos(P[4], P[0]), gm1);
    const R es = e0 + RM::div_pos(R(0.5) * bb, P[0]);
    const R p0 = P[4];
    const R ps = p0 + R(0.5) * bb;
    const R hs = R(1.0) + es + RM::div_pos(ps, P[0]);
    if (vv >= R(1)) return ERR_SUPERLUMINAL;
#pragma unroll
    for (int q = 0; q < 8; ++q) S.P[q] = P[q];
    S.U[0] = P[0] * W0;
    S.U[4] = P[0] * hs * W2 - ps - b0 * b0 - S.U[0];
    S.U[1] = P[0] * hs * W2 * P[1] - b0 * b1;
    S.U[2] = P[0] * hs * W2 * P[2] - b0 * b2;
    S.U[3] = P[0] * hs * W2 * P[3] - b0 * b3;
    S.U[5] = P[5];
    S.U[6] = P[6];
    S.U[7] = P[7];
    S.F[0] = vn * S.U[0];
    S.F[1] = vn * S.U[1] - RM::div_pos(Bn * b1, W0) + ps * dAA[0];
    S.F[2] = vn * S.U[2] - RM::div_pos(Bn * b2, W0) + ps * dAA[1];
    S.F[3] = vn * S.U[3] - RM::div_pos(Bn * b3, W0) + ps * dAA[2];
    S.F[4] = vn * S.U[4] - RM::div_pos(Bn * b0, W0) + ps * vn;
    S.F[5] = vn * S.U[5] - Bn * P[1];
    S.F[6] = vn * S.U[6] - Bn * P[2];
    S.F[7] = vn * S.U[7] - Bn * P[3];
    double Pd[8], nd[3], Ad[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) Pd[q] = double(P[q]);
    nd[0] = double(dAA[0]); nd[1] = double(dAA[1]); nd[2] = double(dAA[2]);
    if (rmhd_eigen(double(gamma), Pd, nd, Ad) != ERR_NONE) return ERR_EIGEN;
#pragma unroll
    for (int q = 0; q < 8; ++q) S.A[q] = R(Ad[q]);
    return ERR_NONE;
}

// Law dispatch.  Law: 0 hydro, 1 MHD, 2 RMHD (include/mara_gpu.h), 3 MHD
// with a positive cooling rate: the MHD fields and an auxiliary ninth one,
// U = P, zero flux and eigenvalue (NewtonianMHD::fromPrimitive with
// numAuxiliary = 1, src/ConservationLaws.cpp:479-485).
template <int Law> struct LawNQ { static constexpr int value = Law == 0 ? 5 : Law == 3 ? 9 : 8; };

template <int Law, typename R, typename A = LawMath<Law, R>>
__device__ __forceinline__ int law_state(R gamma, const R* P, const R* n, State<R, LawNQ<Law>::value>& S)
{
    if constexpr (Law == 0) { hydro_state(gamma, P, n, S); return ERR_NONE; }
    else if constexpr (Law == 1) { mhd_state<R, A>(gamma, P, n, S); return ERR_NONE; }
    else if constexpr (Law == 3)
    {
        State<R, 8> m;
        mhd_state<R, A>(gamma, P, n, m);
#pragma unroll
        for (int q = 0; q < 8; ++q)
        {
            S.P[q] = m.P[q];
            S.U[q] = m.U[q];
            S.F[q] = m.F[q];
            S.A[q] = m.A[q];
        }
        S.P[8] = P[8];
        S.U[8] = P[8];
        S.F[8] = R(0);
        S.A[8] = R(0);
        return ERR_NONE;
    }
    else return rmhd_state(gamma, P, n, S);
}

// ---------------------------------------------------------------------------
// Riemann solvers (src/RiemannSolvers.cpp:38-118)
// ---------------------------------------------------------------------------
// std::max_element / min_element over the reference's 10-slot eigenvalue
// array: the slots beyond NQ hold zeros.
template <typename R, int NQ>
__device__ __forceinline__ R max_elem(const R* A)
{
    R m = A[0];
    if constexpr (NQ == 5)
    {
        // hydro states hold A[1] = A[2] = A[3] = vn (hydro_state): once m has
        // met A[1], comparing it with A[2] and A[3] never changes it
        if (m < A[1]) m = A[1];
        if (m < A[4]) m = A[4];
    }
    else
    {
#pragma unroll
        for (int n = 1; n < NQ; ++n) if (m < A[n]) m = A[n];
    }
    if (m < R(0)) m = R(0);
    return m;
}

template <typename R, int NQ>
__device__ __forceinline__ R min_elem(const R* A)
{
    R m = A[0];
    if constexpr (NQ == 5)
    {
        if (A[1] < m) m = A[1];
        if (A[4] < m) m = A[4];
    }
    else
    {
#pragma unroll
        for (int n = 1; n < NQ; ++n) if (A[n] < m) m = A[n];
    }
    if (R(0) < m) m = R(0);
    return m;
}

template <typename R, int NQ, typename D = Math<R>>
__device__ __forceinline__ void hlle(const State<R, NQ>& L, const State<R, NQ>& Rs, R* F)
{
    R ap, am;
    if constexpr (D::ieee_minmax)
    {
        // the same extremes over both states' eigenvalues and 0 as single
        // min/max instructions (Rel's arithmetic)
        ap = R(0);
        am = R(0);
#pragma unroll
        for (int n = 0; n < NQ; ++n)
        {
            ap = __builtin_fmax(ap, __builtin_fmax(L.A[n], Rs.A[n]));
            am = __builtin_fmin(am, __builtin_fmin(L.A[n], Rs.A[n]));
        }
    }
    else
    {
        const R epl = max_elem<R, NQ>(L.A);
        const R eml = min_elem<R, NQ>(L.A);
        const R epr = max_elem<R, NQ>(Rs.A);
        const R emr = min_elem<R, NQ>(Rs.A);
        ap = max3(R(0.0), epl, epr);
        am = min3(R(0.0), eml, emr);
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        F[q] = D::div_pos(ap * L.F[q] - am * Rs.F[q] + ap * am * (Rs.U[q] - L.U[q]), ap - am);
}

template <typename R>
__device__ __forceinline__ void hllc(const State<R, 5>& L, const State<R, 5>& Rs, const R* dA, R* Fout)
{
    const R epl = max_elem<R, 5>(L.A);
    const R eml = min_elem<R, 5>(L.A);
    const R epr = max_elem<R, 5>(Rs.A);
    const R emr = min_elem<R, 5>(Rs.A);
    const R ap = max3(R(0.0), epl, epr);
    const R am = min3(R(0.0), eml, emr);
    const R* Pl = L.P;
    const R* Pr = Rs.P;
    const R* Ul = L.U;
    const R* Ur = Rs.U;
    const R vnr = Pr[1] * dA[0] + Pr[2] * dA[1] + Pr[3] * dA[2];
    const R vnl = Pl[1] * dA[0] + Pl[2] * dA[1] + Pl[3] * dA[2];
    const R lc = Math<R>::div((Pr[4] - Pr[0] * vnr * (ap - vnr)) - (Pl[4] - Pl[0] * vnl * (am - vnl)),
                              Pl[0] * (am - vnl) - Pr[0] * (ap - vnr));
    // Only the star state on the contact's side of the face enters the flux
    // (left when 0 <= lc); each lane builds that one, with the same
    // expressions as the reference's left and right star states.
    const R s = R(0.0);
    const bool left = s <= lc;
    R Ps[5], Us_side[5], Fs[5];
#pragma unroll
    for (int i = 0; i < 5; ++i)
    {
        Ps[i] = left ? Pl[i] : Pr[i];
        Us_side[i] = left ? Ul[i] : Ur[i];
        Fs[i] = left ? L.F[i] : Rs.F[i];
    }
    const R a = left ? am : ap;
    const R vn = left ? vnl : vnr;
    const R f = Math<R>::div(Ps[0] * (a - vn), a - lc);
    R Ustar[5];
    Ustar[0] = f;
    Ustar[4] = f * (Math<R>::div(Us_side[4], Ps[0]) + (lc - vn) * (lc + Math<R>::div(Ps[4], Ps[0] * (a - vn))));
    Ustar[1] = f * ((lc - vn) * dA[0] + Ps[1]);
    Ustar[2] = f * ((lc - vn) * dA[1] + Ps[2]);
    Ustar[3] = f * ((lc - vn) * dA[2] + Ps[3]);
    if (s <= am)
    {
#pragma unroll
        for (int i = 0; i < 5; ++i) Fout[i] = L.F[i];
    }
    else if ((am < s && s <= lc) || (lc < s && s <= ap))
    {
#pragma unroll
        for (int i = 0; i < 5; ++i) Fout[i] = Fs[i] + a * (Ustar[i] - Us_side[i]);
    }
    else if (ap < s)
    {
#pragma unroll
        for (int i = 0; i < 5; ++i) Fout[i] = Rs.F[i];
    }
    else
    {
#pragma unroll
        for (int i = 0; i < 5; ++i) Fout[i] = R(0);
    }
}

// ---------------------------------------------------------------------------
// Primitive <-> conserved for the zone update (binary64 in both forms)
// ---------------------------------------------------------------------------
// U part of fromPrimitive, as FieldOperator::generateConserved uses it.
template <int Law>
__device__ __forceinline__ int prim_to_cons(double gamma, const double* P, double* U)
{
    if constexpr (Law == 0)
    {
        const double gm1 = gamma - 1.0;
        const double vv = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
        U[0] = P[0];
        U[1] = P[0] * P[1];
        U[2] = P[0] * P[2];
        U[3] = P[0] * P[3];
        U[4] = P[0] * 0.5 * vv + P[4] / gm1;
        return ERR_NONE;
    }
    else if constexpr (Law == 1)
    {
        const double gm1 = gamma - 1.0;
        const double vv = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
        const double BB = P[5] * P[5] + P[6] * P[6] + P[7] * P[7];
        U[0] = P[0];
        U[1] = P[0] * P[1];
        U[2] = P[0] * P[2];
        U[3] = P[0] * P[3];
        U[4] = P[0] * 0.5 * vv + 0.5 * BB + P[4] / gm1;
        U[5] = P[5];
        U[6] = P[6];
        U[7] = P[7];
        return ERR_NONE;
    }
    else if constexpr (Law == 3)
    {
        prim_to_cons<1>(gamma, P, U);
        U[8] = P[8];
        return ERR_NONE;
    }
    else
    {
#pragma clang fp contract(fast)
        const double gm1 = gamma - 1.0;
        const double vv = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
        const double BB = P[5] * P[5] + P[6] * P[6] + P[7] * P[7];
        const double Bv = P[5] * P[1] + P[6] * P[2] + P[7] * P[3];
        const double W2 = Rel::rcp(1.0 - vv);
        const double W0 = Rel::sqrt(W2);
        const double b0 = W0 * Bv;
        const double bb = Rel::div(BB + b0 * b0, W2);
        const double b1 = Rel::div(P[5] + b0 * W0 * P[1], W0);
        const double b2 = Rel::div(P[6] + b0 * W0 * P[2], W0);
        const double b3 = Rel::div(P[7] + b0 * W0 * P[3], W0);
        const double e0 = Rel::div(Rel::div(P[4], P[0]), gm1);
        const double es = e0 + Rel::div(0.5 * bb, P[0]);
        const double p0 = P[4];
        const double ps = p0 + 0.5 * bb;
        const double hs = 1.0 + es + Rel::div(ps, P[0]);
        if (vv >= 1) return ERR_SUPERLUMINAL;
        U[0] = P[0] * W0;
        U[4] = P[0] * hs * W2 - ps - b0 * b0 - U[0];
        U[1] = P[0] * hs * W2 * P[1] - b0 * b1;
        U[2] = P[0] * hs * W2 * P[2] - b0 * b2;
        U[3] = P[0] * hs * W2 * P[3] - b0 * b3;
        U[5] = P[5];
        U[6] = P[6];
        U[7] = P[7];
        return ERR_NONE;
    }
}

// Conserved -> primitive inversion of the relativistic MHD law
// (ConsToPrimRMHD, src/ConsToPrimRMHD.cpp:77-323): a 2-D Newton iteration on
// (Z, W) = (rho h W^2, Lorentz factor) after Anton et al. (2006), and where it
// fails a 1-D Newton iteration on Z after Noble et al. (2006); both stop at a
// relative step below 1e-12 or after 250 iterations.  The operation order is
// the reference's: which zones converge, and after how many iterations, is
// part of the result.
struct Inversion
{
    double k_gas;     // (gamma - 1) / gamma: p = k_gas (rho h - rho)
    double pfloor;    // pressure floor relative to rho, <= 0 off
    double D, tau, S2, B2, BS, BS2;
    double U[8];
    double Z0, W0;    // starting point: Z = sqrt(S^2 + D^2), W = Z / D
};

__device__ __forceinline__ bool finite8(const double* x)
{
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 8; ++i) ok = ok && isfinite(x[i]);
    return ok;
}

// status codes of the inversion (0 success; 1 NaN in U, 2 D < 0, 3 tau < 0,
// 4 NaN in P, 5 p < 0, 6 rho < 0, 7 v.v >= 1, 8 iteration limit)
__device__ __forceinline__ int inversion_check_input(const double* U)
{
    if (U[0] < 0.0) return 2;
    if (U[4] < 0.0) return 3;
    if (!finite8(U)) return 1;
    return 0;
}

// the primitives of a converged (Z, W)
__device__ __forceinline__ int inversion_primitives(const Inversion& s, double Z, double W, double* Pout)
{
#pragma clang fp contract(fast)
    double P[8];
    const double bw = Rel::div(s.BS * W, Z);
    P[0] = Rel::div(s.D, W);
    P[4] = Rel::div(s.D, W) * (Rel::div(Z, s.D * W) - 1.0) * s.k_gas;
    P[1] = Rel::div(s.U[1] + Rel::div(bw * s.U[5], W), Z + s.B2);
    P[2] = Rel::div(s.U[2] + Rel::div(bw * s.U[6], W), Z + s.B2);
    P[3] = Rel::div(s.U[3] + Rel::div(bw * s.U[7], W), Z + s.B2);
    P[5] = s.U[5];
    P[6] = s.U[6];
    P[7] = s.U[7];
    if (P[4] <= 0.0 && s.pfloor > 0.0) P[4] = P[0] * s.pfloor;
    const double v2 = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
    if (v2 >= 1.0) return 7;
    if (P[4] < 0.0) return 5;
    if (P[0] < 0.0) return 6;
    if (!finite8(P)) return 4;
#pragma unroll
    for (int i = 0; i < 8; ++i) Pout[i] = P[i];
    return 0;
}

// Newton on the residuals r_mom (momentum: S^2 as a function of Z, W) and
// r_en (energy: tau) in (Z, W), with the inverse Jacobian written out
__device__ inline int invert_zw(const Inversion& s, double* P)
{
#pragma clang fp contract(fast)
    if (int bad = inversion_check_input(s.U)) return bad;
    int iterations = 0;
    double step = 1.0;
    double W = s.W0, Z = s.Z0;
    const double D = s.D, tau = s.tau, S2 = s.S2, B2 = s.B2, BS2 = s.BS2, kg = s.k_gas;
    while (step > 1e-12)
    {
        const double Z2 = Z * Z;
        const double Z3 = Z * Z2;
        const double W2 = W * W;
        const double W3 = W * W2;
        const double p = Rel::div(D, W) * (Rel::div(Z, D * W) - 1.0) * kg;
        const double J00 = Rel::div(2 * (B2 + Z) * (BS2 * W2 + (W2 - 1) * Z3), W2 * Z3);
        const double J01 = Rel::div(2 * (B2 + Z) * (B2 + Z), W3);
        const double J10 = 1.0 + Rel::div(BS2, Z3) - Rel::div(kg, W2);
        const double J11 = Rel::div(B2, W3) + Rel::div(2 * Z - D * W, W3) * kg;
        const double r_mom = -S2 + Rel::div((Z + B2) * (Z + B2) * (W2 - 1), W2) - Rel::div((2 * Z + B2) * BS2, Z2);
        const double r_en = -tau + Z + B2 - p - Rel::div(0.5 * B2, W2) - Rel::div(0.5 * BS2, Z2) - D;
        const double det = J00 * J11 - J01 * J10;
        const double K00 = Rel::div(J11, det), K01 = Rel::div(-J01, det);
        const double K10 = Rel::div(-J10, det), K11 = Rel::div(J00, det);
        const double dZ = -(K00 * r_mom + K01 * r_en);
        const double dW = -(K10 * r_mom + K11 * r_en);
        double Zn = Z + dZ;
        double Wn = W + dW;
        Zn = (Zn > 0.0) ? Zn : -Zn;
        Zn = (Zn < 1e20) ? Zn : Z;
        Wn = (Wn > 1.0) ? Wn : 1.0;
        Wn = (Wn < 1e12) ? Wn : 1e12;
        Z = Zn;
        W = Wn;
        step = fabs(Rel::div(dZ, Z)) + fabs(Rel::div(dW, W));
        ++iterations;
        if (iterations == 250) return 8;
    }
    return inversion_primitives(s, Z, W, P);
}

// Newton on the energy residual in Z alone, v^2 and W eliminated through the
// momentum relation v^2 = num / den
__device__ inline int invert_z(const Inversion& s, double* P)
{
#pragma clang fp contract(fast)
    if (int bad = inversion_check_input(s.U)) return bad;
    int iterations = 0;
    double step = 1.0;
    double Z = s.Z0;
    const double D = s.D, tau = s.tau, S2 = s.S2, B2 = s.B2, BS2 = s.BS2, kg = s.k_gas;
    while (step > 1e-12)
    {
        const double Z2 = Z * Z;
        const double Z3 = Z * Z2;
        const double num = S2 * Z2 + BS2 * (B2 + 2 * Z);
        const double den = (B2 + Z) * (B2 + Z) * Z2;
        const double dnum = 2 * (S2 * Z + BS2);
        const double dden = 2 * Z * (B2 + Z) * (B2 + 2 * Z);
        const double v2 = Rel::div(num, den);
        const double W2 = Rel::rcp(1.0 - v2);
        const double W = Rel::sqrt(W2);
        const double W3 = W * W2;
        const double p = Rel::div(D, W) * (Rel::div(Z, D * W) - 1.0) * kg;
        const double dv2_dZ = Rel::div(dnum * den - dden * num, den * den);
        const double dp_dZ_atW = Rel::div(kg, W2);
        const double dp_dW_atZ = kg * (Rel::div(D, W2) - Rel::div(2 * Z, W3));
        const double dW_dv2 = 0.5 * W3;
        const double dp_dZ = dp_dW_atZ * dW_dv2 * dv2_dZ + dp_dZ_atW;
        const double r = tau + D - 0.5 * B2 * (1 + v2) + Rel::div(0.5 * BS2, Z2) - Z + p;
        const double dr = -0.5 * B2 * dv2_dZ - Rel::div(BS2, Z3) - 1.0 + dp_dZ;
        const double dZ = Rel::div(-r, dr);
        double Zn = Z + dZ;
        Zn = (Zn > 0.0) ? Zn : -Zn;
        Zn = (Zn < 1e20) ? Zn : Z;
        Z = Zn;
        step = fabs(Rel::div(dZ, Z));
        ++iterations;
        if (iterations == 250) return 8;
    }
    const double Z2 = Z * Z;
    const double num = S2 * Z2 + BS2 * (B2 + 2 * Z);
    const double den = (B2 + Z) * (B2 + Z) * Z2;
    const double v2 = Rel::div(num, den);
    const double W2 = Rel::rcp(1.0 - v2);
    const double W = Rel::sqrt(W2);
    return inversion_primitives(s, Z, W, P);
}

// fromConserved -> primitive (src/ConservationLaws.cpp:238-251, 370-414,
// 627-657).  Returns 0 on success, 1 on inversion failure; floor = 1 when
// the MHD pressure floor was used.
template <int Law>
__device__ __forceinline__ int cons_to_prim(double gamma, double pfloor, const double* U, double* P, int& floor)
{
    floor = 0;
    if constexpr (Law == 0)
    {
        const double gm1 = gamma - 1.0;
        const double pp = U[1] * U[1] + U[2] * U[2] + U[3] * U[3];
        P[0] = U[0];
        P[4] = (U[4] - 0.5 * pp / U[0]) * gm1;
        P[1] = U[1] / U[0];
        P[2] = U[2] / U[0];
        P[3] = U[3] / U[0];
        return 0;
    }
    else if constexpr (Law == 3)
    {
        // the auxiliary field is copied before the bad-state test
        // (src/ConservationLaws.cpp:386-389)
        P[8] = U[8];
        return cons_to_prim<1>(gamma, pfloor, U, P, floor);
    }
    else if constexpr (Law == 1)
    {
        const double gm1 = gamma - 1.0;
        const double pp = U[1] * U[1] + U[2] * U[2] + U[3] * U[3];
        const double BB = U[5] * U[5] + U[6] * U[6] + U[7] * U[7];
        P[0] = U[0];
        P[4] = (U[4] - 0.5 * pp / U[0] - 0.5 * BB) * gm1;
        P[1] = U[1] / U[0];
        P[2] = U[2] / U[0];
        P[3] = U[3] / U[0];
        P[5] = U[5];
        P[6] = U[6];
        P[7] = U[7];
        if (P[4] < 0.0 || P[0] < 0.0 || U[0] < 0.0 || U[4] < 0.0)
        {
            if (P[4] < 0.0 && pfloor > 0.0)
            {
                P[4] = pfloor * P[0];
                floor = 1;
                return 0;
            }
            return 1;
        }
        return 0;
    }
    else
    {
#pragma clang fp contract(fast)
        Inversion s;
        s.k_gas = (gamma - 1.0) / gamma;
        s.pfloor = pfloor;
        s.D = U[0];
        s.tau = U[4];
        s.S2 = U[1] * U[1] + U[2] * U[2] + U[3] * U[3];
        s.B2 = U[5] * U[5] + U[6] * U[6] + U[7] * U[7];
        s.BS = U[5] * U[1] + U[6] * U[2] + U[7] * U[3];
        s.BS2 = s.BS * s.BS;
#pragma unroll
        for (int i = 0; i < 8; ++i) s.U[i] = U[i];
        s.Z0 = Rel::sqrt(s.S2 + s.D * s.D);
        s.W0 = Rel::div(s.Z0, s.D);
        if (invert_zw(s, P) == 0) return 0;
        if (invert_z(s, P) == 0) return 0;
        return 1;
    }
}

// fromConserved of the mixed form for the hydro and MHD laws: binary64 like
// cons_to_prim, with one reciprocal of the density in place of its four
// quotients (the mixed form is held to a tolerance, not to the reference's
// rounding).  Failure and floor semantics as cons_to_prim.  The RMHD law
// keeps cons_to_prim.
template <int Law>
__device__ __forceinline__ int cons_to_prim_mixed(double gamma, double pfloor, const double* U, double* P,
                                                  int& floor)
{
    static_assert(Law == 0 || Law == 1, "hydro and MHD only");
    floor = 0;
    const double gm1 = gamma - 1.0;
    const double ir = 1.0 / U[0];
    const double pp = U[1] * U[1] + U[2] * U[2] + U[3] * U[3];
    P[0] = U[0];
    P[1] = U[1] * ir;
    P[2] = U[2] * ir;
    P[3] = U[3] * ir;
    if constexpr (Law == 0)
    {
        P[4] = (U[4] - 0.5 * pp * ir) * gm1;
        return 0;
    }
    else
    {
        const double BB = U[5] * U[5] + U[6] * U[6] + U[7] * U[7];
        P[4] = (U[4] - 0.5 * pp * ir - 0.5 * BB) * gm1;
        P[5] = U[5];
        P[6] = U[6];
        P[7] = U[7];
        if (P[4] < 0.0 || P[0] < 0.0 || U[0] < 0.0 || U[4] < 0.0)
        {
            if (P[4] < 0.0 && pfloor > 0.0)
            {
                P[4] = pfloor * P[0];
                floor = 1;
                return 0;
            }
            return 1;
        }
        return 0;
    }
}

} // namespace mara
