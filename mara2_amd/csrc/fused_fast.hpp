// fused_fast.hpp -- stage kernel of the fast forms for the Newtonian hydro
// law with PLM on 2-D and 3-D blocks: the fast binary64 form (R = double)
// and the mixed form (R = float: reconstruction, face states and Riemann
// fluxes in binary32, the zone update in binary64).
//
// Same method and the same 2.5-D tile march as fused.hpp (k_fused_plm):
// half-slopes shared by a zone's two faces, plane k with its x-y halo in LDS,
// the zone's own z column in registers, the next tile-row's faces on the
// tile edge.  What differs is the arithmetic, which does not keep the
// reference's rounding:
//   - divisions are products with a reciprocal and square roots come from a
//     reciprocal square root; in binary64 both are the hardware
//     approximations (v_rcp_f64, v_rsq_f64, relative error ~5e-8) refined by
//     one Newton step, within ~4e-15, in binary32 the hardware ones;
//   - multiply-adds contract to FMA (the translation units are compiled with
//     -ffp-contract=on);
//   - the limiter is minmod written with min/max, the face normals are the
//     coordinate axes (the reference's yhat carries a 6e-17 x component that
//     changes nothing above roundoff), the wave-speed extremes are vn -/+ cs;
//   - HLLC forms the conserved state and flux of the contact's side only and
//     needs one reciprocal for its star state, HLLE one per face;
//   - the zone update takes one reciprocal of the density;
//   - the RK stage state is carried as the primitives alone: a stage's U is
//     U(Pin) of the zone's own binary64 input value (where the reference
//     carries U between stages; the two differ by the rounding of one
//     P(U) -> U(P) round trip), so the kernel neither reads nor writes the U
//     array, and the binary64 form's first stage takes U0 from that value
//     instead of re-reading P0;
//   - arrays read once (P0) or written once (Pout) use non-temporal access,
//     which leaves the L2 to the input planes that neighbouring tiles re-read
//     as their halo.
// The binary64 form is held to relative L1 <= 1e-12 of the reference's
// conserved fields after 10 steps, the mixed form to 1e-5.
#pragma once

#include "fused.hpp"

namespace mara {
namespace fast64 {

template <typename R> struct Fast;

template <> struct Fast<double>
{
    // 1/x: v_rcp_f64 (relative error up to 4.6e-8) and one Newton step, which
    // leaves at most 2.1e-15 (measured on gfx950 over 2^20 log-uniform inputs)
    static __device__ __forceinline__ double rcp(double x)
    {
        const double y = __builtin_amdgcn_rcp(x);
        return __builtin_fma(y, __builtin_fma(-x, y, 1.0), y);
    }
    // 1/sqrt(s) for s > 0: v_rsq_f64 (up to 5.2e-8) and one Newton step
    // y += y (1 - s y^2) / 2, at most 4.1e-15 after it (measured likewise)
    static __device__ __forceinline__ double rsqrt(double s)
    {
        const double r = __builtin_amdgcn_rsq(s);
        return __builtin_fma(0.5 * r, __builtin_fma(-(s * r), r, 1.0), r);
    }
};

template <> struct Fast<float>
{
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float rsqrt(float s) { return __builtin_amdgcn_rsqf(s); }
};

// sqrt(x / y) for x >= 0, y > 0 as x / sqrt(x y); 0 when x == 0
template <typename R>
__device__ __forceinline__ R sqrt_ratio(R x, R y)
{
    const R s = x * y;
    return s == R(0) ? R(0) : x * Fast<R>::rsqrt(s);
}

// 0.5 * minmod(theta (v0 - vl), (vr - vl) / 2, theta (vr - v0)): the value
// of least magnitude when all three share a sign, else 0 (th2 = theta / 2):
// max(min(a, b, c), min(max(a, b, c), 0))
template <typename R>
__device__ __forceinline__ R half_slope(R th2, R vl, R v0, R vr)
{
    const R dl = v0 - vl, dr = vr - v0;
    const R a = th2 * dl, c = th2 * dr, b = R(0.25) * (dl + dr);
    const R lo = __builtin_fmin(a, __builtin_fmin(b, c));
    const R hi = __builtin_fmax(a, __builtin_fmax(b, c));
    return __builtin_fmax(lo, __builtin_fmin(hi, R(0)));
}

template <typename R>
struct Consts
{
    R gamma, igm1, th2;
};

// sound speed sqrt(gamma p / rho)
template <typename R>
__device__ __forceinline__ R sound(const Consts<R>& K, R rho, R p) { return sqrt_ratio(K.gamma * p, rho); }

// HLLC flux (Toro eqs. 10.33, 10.58; HllcNewtonianHydroRiemannSolver,
// src/RiemannSolvers.cpp:63-118) through a face normal to axis AX from the
// primitive states L and R.  The signal speeds are the extremes of the two
// states' eigenvalues and 0 (am <= 0 <= ap); the flux is the contact side's
// F + a (U* - U), which is that side's F where the side is supersonic (a = 0).
template <int AX, typename R>
__device__ __forceinline__ void hllc(const Consts<R>& K, const R* L, const R* Rt, R* F)
{
    const R vnl = L[1 + AX], vnr = Rt[1 + AX];
    const R csl = sound(K, L[0], L[4]), csr = sound(K, Rt[0], Rt[4]);
    const R am = __builtin_fmin(R(0), __builtin_fmin(vnl - csl, vnr - csr));
    const R ap = __builtin_fmax(R(0), __builtin_fmax(vnl + csl, vnr + csr));
    const R dl = am - vnl, dr = ap - vnr;
    const R lc = ((Rt[4] - Rt[0] * vnr * dr) - (L[4] - L[0] * vnl * dl)) * Fast<R>::rcp(L[0] * dl - Rt[0] * dr);
    const bool left = R(0) <= lc;
    R P[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) P[q] = left ? L[q] : Rt[q];
    const R a = left ? am : ap;
    const R D = left ? dl : dr;  // a - vn
    const R r = P[0], p = P[4], vn = P[1 + AX];
    const R E = r * R(0.5) * (P[1] * P[1] + P[2] * P[2] + P[3] * P[3]) + p * K.igm1;
    const R k = a - lc;
    const R ik = k != R(0) ? Fast<R>::rcp(k) : R(0);
    const R f = r * D * ik;  // rho* = rho (a - vn) / (a - lc)
    F[0] = vn * r + a * (f - r);
#pragma unroll
    for (int t = 0; t < 3; ++t)
    {
        const R Ut = r * P[1 + t];
        const R Us = f * (t == AX ? lc : P[1 + t]);
        F[1 + t] = vn * Ut + (t == AX ? p : R(0)) + a * (Us - Ut);
    }
    const R Es = ik * (D * E + (lc - vn) * (r * D * lc + p));
    F[4] = vn * (E + p) + a * (Es - E);
}

// HLLE flux (HlleRiemannSolver, src/RiemannSolvers.cpp:38-57)
template <int AX, typename R>
__device__ __forceinline__ void hlle(const Consts<R>& K, const R* L, const R* Rt, R* F)
{
    const R vnl = L[1 + AX], vnr = Rt[1 + AX];
    const R csl = sound(K, L[0], L[4]), csr = sound(K, Rt[0], Rt[4]);
    const R am = __builtin_fmin(R(0), __builtin_fmin(vnl - csl, vnr - csr));
    const R ap = __builtin_fmax(R(0), __builtin_fmax(vnl + csl, vnr + csr));
    const R iw = Fast<R>::rcp(ap - am);
    const R El = L[0] * R(0.5) * (L[1] * L[1] + L[2] * L[2] + L[3] * L[3]) + L[4] * K.igm1;
    const R Er = Rt[0] * R(0.5) * (Rt[1] * Rt[1] + Rt[2] * Rt[2] + Rt[3] * Rt[3]) + Rt[4] * K.igm1;
    R Ul[5], Ur[5], Fl[5], Fr[5];
    Ul[0] = L[0];
    Ur[0] = Rt[0];
#pragma unroll
    for (int t = 0; t < 3; ++t)
    {
        Ul[1 + t] = L[0] * L[1 + t];
        Ur[1 + t] = Rt[0] * Rt[1 + t];
    }
    Ul[4] = El;
    Ur[4] = Er;
#pragma unroll
    for (int q = 0; q < 4; ++q)
    {
        Fl[q] = vnl * Ul[q];
        Fr[q] = vnr * Ur[q];
    }
    Fl[1 + AX] += L[4];
    Fr[1 + AX] += Rt[4];
    Fl[4] = vnl * (El + L[4]);
    Fr[4] = vnr * (Er + Rt[4]);
    const R apm = ap * am;
#pragma unroll
    for (int q = 0; q < 5; ++q) F[q] = (ap * Fl[q] - am * Fr[q] + apm * (Ur[q] - Ul[q])) * iw;
}

template <int Riem, int AX, typename R>
__device__ __forceinline__ void face_flux(const Consts<R>& K, const R* L, const R* Rt, R* F)
{
    if constexpr (Riem == 1) hllc<AX>(K, L, Rt, F);
    else hlle<AX>(K, L, Rt, F);
}

// U(P) of NewtonianHydro with 1 / (gamma - 1) as a product, binary64
__device__ __forceinline__ void cons_of(double igm1, const double* P, double* U)
{
    U[0] = P[0];
    U[1] = P[0] * P[1];
    U[2] = P[0] * P[2];
    U[3] = P[0] * P[3];
    U[4] = P[0] * 0.5 * (P[1] * P[1] + P[2] * P[2] + P[3] * P[3]) + P[4] * igm1;
}

// P(U) with one reciprocal of the density, binary64
__device__ __forceinline__ void prim_of(double gm1, const double* U, double* P)
{
    const double ir = Fast<double>::rcp(U[0]);
    P[0] = U[0];
    P[1] = U[1] * ir;
    P[2] = U[2] * ir;
    P[3] = U[3] * ir;
    P[4] = (U[4] - 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) * ir) * gm1;
}

// CFL time step of a zone: Lmin / max over axes of (|v_axis| + cs), the
// reference's min over zones of L / max |eigenvalue| with axis normals,
// in the face arithmetic
template <typename R>
__device__ __forceinline__ double courant(const Consts<R>& K, const double* Pd, double Lmin)
{
    const R P[5] = {R(Pd[0]), R(Pd[1]), R(Pd[2]), R(Pd[3]), R(Pd[4])};
    const R cs = sound(K, P[0], P[4]);
    const R vm = __builtin_fmax(__builtin_fabs(P[1]), __builtin_fmax(__builtin_fabs(P[2]), __builtin_fabs(P[3])));
    return double(R(Lmin) * Fast<R>::rcp(vm + cs));
}

// Tiles of 64 x TY zones marching over z planes: binary64 64 x 4 (three
// work-groups per CU within the LDS), binary32 64 x 8.
template <typename R> struct FastTile
{
    static constexpr int TX = 64;
    static constexpr int TY = sizeof(R) == 8 ? 4 : 8;
    static constexpr int WAVES = sizeof(R) == 8 ? 3 : 4;
};

// One tile (tx, ty, tz) of the stage; every thread of the work-group calls
// it (it synchronises internally).
template <typename R, int Riem, bool Z>
__device__ __forceinline__ void fast_plm_tile(const double* __restrict__ Pin, const double* __restrict__ P0,
                                              double* __restrict__ Pout, const Geom& G, double b, int first_stage,
                                              int courant_out, Ctl* ctl, int tx, int ty, int tz, int kz,
                                              bool wrap)
{
    constexpr bool F64 = sizeof(R) == 8;
    constexpr int NG = 2, TX = FastTile<R>::TX, TY = FastTile<R>::TY;
    constexpr int HX = TX + 2 * NG, HY = TY + 2 * NG;
    constexpr int NT = TX * TY;
    __shared__ R plane[5][HY][HX];   // plane k, halo 2; zone (i, j) at [j - j0 + 2][i - i0 + 2]
    __shared__ R sy[5][TY + 2][TX];  // y half-slopes of rows -1 .. TY at [r + 1]
    __shared__ R sxe[5][TY][3];      // x half-slopes of zones i0 - 1, i0 + 63, i0 + 64
    __shared__ R fy[5][TY + 1][TX];  // y-face fluxes, face j0 + r at [r]
    __shared__ R fxe[5][TY];         // x-face i0 + 64 of each row
    __shared__ double wmin[NT / 64];

    Consts<R> K;
    K.gamma = R(G.gamma);
    K.igm1 = R(1.0 / (G.gamma - 1.0));
    K.th2 = R(0.5 * G.theta);
    const double gm1 = G.gamma - 1.0, igm1 = 1.0 / gm1;
    double dmin = 1.79769313486231570815e+308;
    const int lane = threadIdx.x;
    const int row = threadIdx.y;
    const int tid = row * TX + lane;
    const int i0 = G.g[0] + tx * TX;
    const int j0 = G.g[1] + ty * TY;
    const int i = i0 + lane, j = j0 + row;
    const int ie = G.N[0] - G.g[0], je = G.N[1] - G.g[1];
    const bool cell_xy = i < ie && j < je;
    const bool xface = i <= ie && j < je;
    const bool yface = j <= je && i < ie;
    const int kb = Z ? G.g[2] + tz * kz : 0;
    const int ke = Z ? min(kb + kz, G.N[2] - G.g[2]) : 1;
    const long long sxy = (long long)G.N[0] * G.N[1];
    const R A0 = R(G.dA[0]), A1 = R(G.dA[1]), A2 = R(G.dA[2]);
    const R fac = R(-1.0 / G.dV);
    const double dt = ctl->dt;
    // wrap (periodic blocks): the halo, the zone columns past the tile's
    // edges and the z window's guard planes come from their interior images,
    // so the stage reads no guard zone (periodic_index)
    const int ic = wrap ? periodic_index(G, 0, i) : min(i, G.N[0] - 1);
    const int jc = wrap ? periodic_index(G, 1, j) : min(j, G.N[1] - 1);
    const long long col = (long long)jc * G.N[0] + ic;
    const auto zplane = [&](int kk) { return (long long)(wrap ? periodic_index(G, 2, kk) : kk) * sxy; };

    R W[5][4];  // z window P(k - 1 .. k + 2) of the zone's column
    R hzc[5];   // z half-slope of zone k
    R Fzlo[5];
    if constexpr (Z)
    {
        R w0[5];
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            w0[q] = R(Pin[q * G.fs + col + zplane(kb - 2)]);
#pragma unroll
            for (int m = 0; m < 3; ++m) W[q][m + 1] = R(Pin[q * G.fs + col + zplane(kb - 1 + m)]);
        }
        R PL[5], PR[5];
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            const R hlo = half_slope(K.th2, w0[q], W[q][1], W[q][2]);
            hzc[q] = half_slope(K.th2, W[q][1], W[q][2], W[q][3]);
            PL[q] = W[q][1] + hlo;
            PR[q] = W[q][2] - hzc[q];
        }
        face_flux<Riem, 2>(K, PL, PR, Fzlo);
    }

    for (int k = kb; k < ke; ++k)
    {
        const long long pbase = (long long)k * sxy;
        // the zone's offset in the plane, opaque per plane like the ring's
        int colk = (int)col;
        asm volatile("" : "+v"(colk));
        // plane k: the ring from memory (PlaneRing) and (3-D) the own zone
        // from the z window; the window's next value is loaded with the ring
        PlaneRing<R, TX, TY, Z> ring;
        ring.load(Pin, G, pbase, i0, j0, tid, wrap);
        R wn[5];
        if constexpr (Z)
        {
#pragma unroll
            for (int q = 0; q < 5; ++q) wn[q] = R(Pin[q * G.fs + colk + zplane(k + 2)]);
#pragma unroll
            for (int q = 0; q < 5; ++q) plane[q][row + NG][lane + NG] = W[q][2];
        }
        ring.store(plane);
        R Fzhi[5];
        if constexpr (Z)
        {
            R PL[5], PR[5];
#pragma unroll
            for (int q = 0; q < 5; ++q)
            {
                W[q][0] = W[q][1];
                W[q][1] = W[q][2];
                W[q][2] = W[q][3];
                W[q][3] = wn[q];
                const R hn = half_slope(K.th2, W[q][1], W[q][2], W[q][3]);
                PL[q] = W[q][1] + hzc[q];
                PR[q] = W[q][2] - hn;
                hzc[q] = hn;
            }
            if (cell_xy) face_flux<Riem, 2>(K, PL, PR, Fzhi);
        }
        __syncthreads();

        // half-slopes: own zone in x (register) and y (LDS); tile-edge zones
        R hx[5];
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            hx[q] = half_slope(K.th2, plane[q][row + 2][lane + 1], plane[q][row + 2][lane + 2],
                               plane[q][row + 2][lane + 3]);
            sy[q][row + 1][lane] = half_slope(K.th2, plane[q][row + 1][lane + 2], plane[q][row + 2][lane + 2],
                                              plane[q][row + 3][lane + 2]);
        }
        if (lane == TX - 1)
        {
#pragma unroll
            for (int q = 0; q < 5; ++q) sxe[q][row][1] = hx[q];
        }
        if (tid < 2 * TX + 2 * TY)
        {
            if (tid < 2 * TX)
            {
                const int cx = tid % TX;
                const int r = tid < TX ? -1 : TY;
#pragma unroll
                for (int q = 0; q < 5; ++q)
                    sy[q][r + 1][cx] = half_slope(K.th2, plane[q][r + 1][cx + 2], plane[q][r + 2][cx + 2],
                                                  plane[q][r + 3][cx + 2]);
            }
            else
            {
                const int e = tid - 2 * TX;
                const int r = e % TY;
                const int xs = e < TY ? 1 : TX + 2;
#pragma unroll
                for (int q = 0; q < 5; ++q)
                    sxe[q][r][e < TY ? 0 : 2] = half_slope(K.th2, plane[q][r + 2][xs - 1], plane[q][r + 2][xs],
                                                           plane[q][r + 2][xs + 1]);
            }
        }
        __syncthreads();

        R Fx[5], Fy[5];
        R hl[5];
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            hl[q] = shift_up(hx[q], sxe[q][row][0]);
        }
        if (xface)
        {
            R PL[5], PR[5];
#pragma unroll
            for (int q = 0; q < 5; ++q)
            {
                PL[q] = plane[q][row + 2][lane + 1] + hl[q];
                PR[q] = plane[q][row + 2][lane + 2] - hx[q];
            }
            face_flux<Riem, 0>(K, PL, PR, Fx);
        }
        if (yface)
        {
            R PL[5], PR[5];
#pragma unroll
            for (int q = 0; q < 5; ++q)
            {
                PL[q] = plane[q][row + 1][lane + 2] + sy[q][row][lane];
                PR[q] = plane[q][row + 2][lane + 2] - sy[q][row + 1][lane];
            }
            face_flux<Riem, 1>(K, PL, PR, Fy);
#pragma unroll
            for (int q = 0; q < 5; ++q) fy[q][row][lane] = Fy[q];
        }
        const int te = tid - (NT - TX - TY);
        if (te >= 0)
        {
            R PL[5], PR[5], F[5];
            if (te < TY)
            {
                const int r = te;
                if (i0 + TX <= ie && j0 + r < je)
                {
#pragma unroll
                    for (int q = 0; q < 5; ++q)
                    {
                        PL[q] = plane[q][r + 2][TX + 1] + sxe[q][r][1];
                        PR[q] = plane[q][r + 2][TX + 2] - sxe[q][r][2];
                    }
                    face_flux<Riem, 0>(K, PL, PR, F);
#pragma unroll
                    for (int q = 0; q < 5; ++q) fxe[q][r] = F[q];
                }
            }
            else
            {
                const int cx = te - TY;
                if (j0 + TY <= je && i0 + cx < ie)
                {
#pragma unroll
                    for (int q = 0; q < 5; ++q)
                    {
                        PL[q] = plane[q][TY + 1][cx + 2] + sy[q][TY][cx];
                        PR[q] = plane[q][TY + 2][cx + 2] - sy[q][TY + 1][cx];
                    }
                    face_flux<Riem, 1>(K, PL, PR, F);
#pragma unroll
                    for (int q = 0; q < 5; ++q) fy[q][TY][cx] = F[q];
                }
            }
        }
        __syncthreads();

        R Fxr[5];
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            Fxr[q] = shift_down(Fx[q], fxe[q][row]);
        }
        if (cell_xy)
        {
            const long long c = pbase + colk;
            // the stage's own state U(Pin) of the zone's binary64 input value
            // (in binary64 the z window or the LDS plane hold it; the binary32
            // tiles re-read it); the step's start state U(P0) is that value
            // again in the first stage, else read from P0
            double pin[5], Us[5], U0[5];
#pragma unroll
            for (int q = 0; q < 5; ++q)
            {
                if constexpr (F64) pin[q] = Z ? W[q][1] : plane[q][row + 2][lane + 2];
                else pin[q] = __builtin_nontemporal_load(&Pin[q * G.fs + c]);
            }
            cons_of(igm1, pin, Us);
            if (first_stage)
            {
#pragma unroll
                for (int q = 0; q < 5; ++q) U0[q] = Us[q];
            }
            else
            {
                double p0[5];
#pragma unroll
                for (int q = 0; q < 5; ++q) p0[q] = __builtin_nontemporal_load(&P0[q * G.fs + c]);
                cons_of(igm1, p0, U0);
            }
            double Un[5];
#pragma unroll
            for (int q = 0; q < 5; ++q)
            {
                R acc = (Fxr[q] - Fx[q]) * A0 + (fy[q][row + 1][lane] - Fy[q]) * A1;
                if constexpr (Z) acc += (Fzhi[q] - Fzlo[q]) * A2;
                const double L = double(fac * acc);
                Un[q] = U0[q] * (1 - b) + (Us[q] + dt * L) * b;
            }
            double Pn[5];
            prim_of(gm1, Un, Pn);
#pragma unroll
            for (int q = 0; q < 5; ++q) __builtin_nontemporal_store(Pn[q], &Pout[q * G.fs + c]);
            if (courant_out)
            {
                const double d = courant(K, Pn, G.Lmin);
                if (d < dmin) dmin = d;
            }
        }
        if constexpr (Z)
        {
#pragma unroll
            for (int q = 0; q < 5; ++q) Fzlo[q] = Fzhi[q];
        }
        // In 3-D the update reads only registers and fy / fxe, which the next
        // plane rewrites after two more barriers every wave must reach, so the
        // next plane's staging need not wait here; the 2-D update reads the
        // plane itself.
        if constexpr (!Z) __syncthreads();
    }
    if (courant_out) courant_to_bins(dmin, wmin, NT / 64, tid, ctl);
}

template <typename R, int Riem, bool Z>
static __global__ void __launch_bounds__(FastTile<R>::TX * FastTile<R>::TY)
    __attribute__((amdgpu_waves_per_eu(FastTile<R>::WAVES, 8)))
k_fast_plm(const double* __restrict__ Pin, const double* __restrict__ P0, double* __restrict__ Pout, Geom G,
           double b, int first_stage, int courant_out, Ctl* ctl, int kz, int wrap)
{
    if (!ctl->active) return;
    constexpr int TX = FastTile<R>::TX, TY = FastTile<R>::TY;
    const int ntx = (G.n[0] + TX - 1) / TX, nty = (G.n[1] + TY - 1) / TY;
    const int ntz = Z ? (G.n[2] + kz - 1) / kz : 1;
    const int tile = xcd_tile(blockIdx.x, ntx * nty * ntz);
    const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
    fast_plm_tile<R, Riem, Z>(Pin, P0, Pout, G, b, first_stage, courant_out, ctl, tx, ty, tz, kz, wrap != 0);
}

// the blocks this kernel serves: hydro, PLM, 2-D or 3-D, two or three stages
inline bool supported(const mara_desc* d)
{
    return d->law == MARA_LAW_HYDRO && d->recon == MARA_RECON_PLM && d->cells[1] > 1 && d->rk_order >= 2
           && d->geometry == MARA_GEOM_CARTESIAN;
}

// Work-groups of the 3-D fast kernel resident on one CU and its tiles per z
// layer (fused_round_t's counterpart).
template <typename R>
inline void tile_round(const Geom& G, int riemann, int* per_cu, int* per_layer)
{
    constexpr int TX = FastTile<R>::TX, TY = FastTile<R>::TY;
    *per_layer = ((G.n[0] + TX - 1) / TX) * ((G.n[1] + TY - 1) / TY);
    int n = 0;
    const hipError_t e = riemann == MARA_RIEMANN_HLLC
                             ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fast_plm<R, 1, true>, TX * TY, 0)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fast_plm<R, 0, true>, TX * TY, 0);
    (void)hipGetLastError();
    *per_cu = e == hipSuccess ? n : 0;
}

// One stage on the block's stream: R = double for the fast binary64 form,
// float for the mixed form.
template <typename R>
inline void launch(const Geom& G, hipStream_t stream, int riemann, const double* Pin, const double* P0,
                   double* Pout, double b, int first, int cout, Ctl* ctl, int kz, int wrap)
{
    constexpr int TX = FastTile<R>::TX, TY = FastTile<R>::TY;
    const bool z = G.N[2] > 1;
    const bool hllc = riemann == MARA_RIEMANN_HLLC;
    const dim3 block(TX, TY, 1);
    const dim3 grid(((G.n[0] + TX - 1) / TX) * ((G.n[1] + TY - 1) / TY) * (z ? (G.n[2] + kz - 1) / kz : 1));
    if (z)
    {
        if (hllc) k_fast_plm<R, 1, true><<<grid, block, 0, stream>>>(Pin, P0, Pout, G, b, first, cout, ctl, kz, wrap);
        else k_fast_plm<R, 0, true><<<grid, block, 0, stream>>>(Pin, P0, Pout, G, b, first, cout, ctl, kz, wrap);
    }
    else
    {
        if (hllc) k_fast_plm<R, 1, false><<<grid, block, 0, stream>>>(Pin, P0, Pout, G, b, first, cout, ctl, kz, wrap);
        else k_fast_plm<R, 0, false><<<grid, block, 0, stream>>>(Pin, P0, Pout, G, b, first, cout, ctl, kz, wrap);
    }
}

// ---------------------------------------------------------------------------
// The CFL time step in the arithmetic of the form whose stage kernels produce
// it in their last stage, so the first step of a run (or of a restart from a
// checkpoint) takes the same dt a later step would take for the same state:
// the reference's binary64 (k_courant), binary32 wave speeds (the mixed
// flux-array, pencil, 2-D MHD and non-PLM fused kernels), the fast tiles'
// courant<double> / courant<float>, or the fast binary64 2-D MHD tiles'
// wave speeds in Rel arithmetic.
// ---------------------------------------------------------------------------
enum CflArith { CFL_EXACT = 0, CFL_MIXED = 1, CFL_FAST64 = 2, CFL_FAST32 = 3, CFL_REL64 = 4 };

template <int Law>
__device__ __forceinline__ double zone_courant_form(int arith, const double* __restrict__ P, const Geom& G,
                                                    long long c, double Lz, int& err)
{
    if constexpr (Law == 0)
    {
        if (arith == CFL_FAST64 || arith == CFL_FAST32)
        {
            double p[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) p[q] = P[q * G.fs + c];
            if (arith == CFL_FAST64)
            {
                Consts<double> K;
                K.gamma = G.gamma;
                return courant(K, p, G.Lmin);
            }
            Consts<float> K;
            K.gamma = float(G.gamma);
            return courant(K, p, G.Lmin);
        }
    }
    if constexpr (Law < 2)
    {
        // the fast binary64 form's 2-D MHD tiles: wave speeds in Rel
        if (arith == CFL_REL64)
        {
            constexpr int NQ = LawNQ<Law>::value;
            double p[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) p[q] = P[q * G.fs + c];
            return courant_of<Law, double, Rel>(p, G, Lz, err);
        }
        if (arith == CFL_MIXED)
        {
            constexpr int NQ = LawNQ<Law>::value;
            double p[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) p[q] = P[q * G.fs + c];
            return courant_of<Law, float>(p, G, Lz, err);
        }
    }
    return zone_courant<Law>(P, G, c, Lz, err);
}

// k_courant with the form's arithmetic
template <int Law>
static __global__ void __launch_bounds__(256) k_courant_form(const double* __restrict__ P, Geom G, Ctl* ctl, int arith)
{
    __shared__ double wmin[4];
    if (!ctl->active) return;
    const long long total = (long long)G.n[0] * G.n[1] * G.n[2];
    const long long nxy = (long long)G.n[0] * G.n[1];
    double dtc = 1.79769313486231570815e+308;
    int err = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x)
    {
        const long long kk = t / nxy;
        const long long r = t - kk * nxy;
        const int j = (int)(r / G.n[0]);
        const int i = (int)(r - (long long)j * G.n[0]);
        const long long c = cell_index(G, G.g[0] + i, G.g[1] + j, G.g[2] + (int)kk);
        const double d =
            zone_courant_form<Law>(arith, P, G, c, zone_length(G, G.g[0] + i, G.g[1] + j, G.g[2] + (int)kk), err);
        if (d < dtc) dtc = d;
    }
    if (err & ERR_SUPERLUMINAL) atomicAdd(&ctl->superluminal, 1ull);
    if (err & ERR_EIGEN) atomicAdd(&ctl->eigen_fail, 1ull);
    courant_to_bins(dtc, wmin, 4, threadIdx.x, ctl);
}

} // namespace fast64
} // namespace mara
