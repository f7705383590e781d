// fused_fast.hpp -- the binary64 stage kernel of the "fast" FP64 form for
// the Newtonian hydro law with PLM on 2-D and 3-D blocks.
//
// Same method and the same 2.5-D tile march as fused.hpp (k_fused_plm):
// half-slopes shared by a zone's two faces, plane k with its x-y halo in LDS,
// the zone's own z column in registers, the next tile-row's faces on the
// tile edge.  What differs is the arithmetic, which is binary64 throughout
// but not the reference's rounding:
//   - divisions are products with a reciprocal (hardware v_rcp_f64 refined by
//     two Newton steps) and square roots come from v_rsq_f64 refined the same
//     way: both within a few ulp, neither correctly rounded;
//   - multiply-adds contract to FMA (the translation unit is compiled with
//     -ffp-contract=fast);
//   - the limiter is minmod written with min/max, the face normals are the
//     coordinate axes (the reference's yhat carries a 6e-17 x component that
//     changes nothing above roundoff), the wave-speed extremes are vn -/+ cs;
//   - HLLC forms the conserved state and flux of the contact's side only and
//     needs one reciprocal for its star state, HLLE one per face;
//   - the zone update takes one reciprocal of the density;
//   - the RK stage state is carried as the primitives alone: a stage's U is
//     U(Pin) of the zone's own input value (where the reference carries U
//     between stages; the two differ by the rounding of one P(U) -> U(P)
//     round trip), so the kernel neither reads nor writes the U array, and
//     the first stage takes U0 from that value instead of re-reading P0.
// The form is held to relative L1 <= 1e-12 of the conserved fields against
// the reference step after 10 steps, not to bit equality.
#pragma once

#include "fused.hpp"

namespace mara {
namespace fast64 {

// 1/x: v_rcp_f64 and two Newton steps
__device__ __forceinline__ double rcp(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-x, y, 1.0);
    return __builtin_fma(y, e, y);
}

// sqrt(x / y) for x >= 0, y > 0 as x * rsqrt(x y): v_rsq_f64 (relative
// error about 2^-23) and two Newton steps y <- y + y (1 - s y^2) / 2; 0 when
// x == 0
__device__ __forceinline__ double sqrt_ratio(double x, double y)
{
    const double s = x * y;
    double r = __builtin_amdgcn_rsq(s);
    double e = __builtin_fma(-(s * r), r, 1.0);
    r = __builtin_fma(0.5 * r, e, r);
    e = __builtin_fma(-(s * r), r, 1.0);
    r = __builtin_fma(0.5 * r, e, r);
    return s == 0.0 ? 0.0 : x * r;
}

// 0.5 * minmod(theta (v0 - vl), (vr - vl) / 2, theta (vr - v0)): the value of
// least magnitude when all three share a sign, else 0 (th2 = theta / 2)
__device__ __forceinline__ double half_slope(double th2, double vl, double v0, double vr)
{
    const double dl = v0 - vl, dr = vr - v0;
    const double a = th2 * dl, c = th2 * dr, b = 0.25 * (dl + dr);
    const double lo = __builtin_fmin(a, __builtin_fmin(b, c));
    const double hi = __builtin_fmax(a, __builtin_fmax(b, c));
    return __builtin_fmax(lo, 0.0) + __builtin_fmin(hi, 0.0);
}

struct Consts
{
    double gamma, igm1, gm1, th2;
};

// sound speed sqrt(gamma p / rho)
__device__ __forceinline__ double sound(const Consts& K, double rho, double p) { return sqrt_ratio(K.gamma * p, rho); }

// HLLC flux (Toro eqs. 10.33, 10.58; HllcNewtonianHydroRiemannSolver,
// src/RiemannSolvers.cpp:63-118) through a face normal to axis AX from the
// primitive states L and R.  The signal speeds are the extremes of the two
// states' eigenvalues and 0 (am <= 0 <= ap); the flux is the contact side's
// F + a (U* - U), which is that side's F where the side is supersonic (a = 0).
template <int AX>
__device__ __forceinline__ void hllc(const Consts& K, const double* L, const double* R, double* F)
{
    const double vnl = L[1 + AX], vnr = R[1 + AX];
    const double csl = sound(K, L[0], L[4]), csr = sound(K, R[0], R[4]);
    const double am = __builtin_fmin(0.0, __builtin_fmin(vnl - csl, vnr - csr));
    const double ap = __builtin_fmax(0.0, __builtin_fmax(vnl + csl, vnr + csr));
    const double dl = am - vnl, dr = ap - vnr;
    const double lc = ((R[4] - R[0] * vnr * dr) - (L[4] - L[0] * vnl * dl)) * rcp(L[0] * dl - R[0] * dr);
    const bool left = 0.0 <= lc;
    double P[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) P[q] = left ? L[q] : R[q];
    const double a = left ? am : ap;
    const double D = left ? dl : dr;  // a - vn
    const double r = P[0], p = P[4], vn = P[1 + AX];
    const double E = r * 0.5 * (P[1] * P[1] + P[2] * P[2] + P[3] * P[3]) + p * K.igm1;
    const double k = a - lc;
    const double ik = k != 0.0 ? rcp(k) : 0.0;
    const double f = r * D * ik;  // rho* = rho (a - vn) / (a - lc)
    F[0] = vn * r + a * (f - r);
#pragma unroll
    for (int t = 0; t < 3; ++t)
    {
        const double Ut = r * P[1 + t];
        const double Us = f * (t == AX ? lc : P[1 + t]);
        F[1 + t] = vn * Ut + (t == AX ? p : 0.0) + a * (Us - Ut);
    }
    const double Es = ik * (D * E + (lc - vn) * (r * D * lc + p));
    F[4] = vn * (E + p) + a * (Es - E);
}

// HLLE flux (HlleRiemannSolver, src/RiemannSolvers.cpp:38-57)
template <int AX>
__device__ __forceinline__ void hlle(const Consts& K, const double* L, const double* R, double* F)
{
    const double vnl = L[1 + AX], vnr = R[1 + AX];
    const double csl = sound(K, L[0], L[4]), csr = sound(K, R[0], R[4]);
    const double am = __builtin_fmin(0.0, __builtin_fmin(vnl - csl, vnr - csr));
    const double ap = __builtin_fmax(0.0, __builtin_fmax(vnl + csl, vnr + csr));
    const double iw = rcp(ap - am);
    const double El = L[0] * 0.5 * (L[1] * L[1] + L[2] * L[2] + L[3] * L[3]) + L[4] * K.igm1;
    const double Er = R[0] * 0.5 * (R[1] * R[1] + R[2] * R[2] + R[3] * R[3]) + R[4] * K.igm1;
    double Ul[5], Ur[5], Fl[5], Fr[5];
    Ul[0] = L[0];
    Ur[0] = R[0];
#pragma unroll
    for (int t = 0; t < 3; ++t)
    {
        Ul[1 + t] = L[0] * L[1 + t];
        Ur[1 + t] = R[0] * R[1 + t];
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
    Fr[1 + AX] += R[4];
    Fl[4] = vnl * (El + L[4]);
    Fr[4] = vnr * (Er + R[4]);
    const double apm = ap * am;
#pragma unroll
    for (int q = 0; q < 5; ++q) F[q] = (ap * Fl[q] - am * Fr[q] + apm * (Ur[q] - Ul[q])) * iw;
}

template <int Riem, int AX>
__device__ __forceinline__ void face_flux(const Consts& K, const double* L, const double* R, double* F)
{
    if constexpr (Riem == 1) hllc<AX>(K, L, R, F);
    else hlle<AX>(K, L, R, F);
}

// U(P) of NewtonianHydro with 1 / (gamma - 1) as a product
__device__ __forceinline__ void cons_of(const Consts& K, const double* P, double* U)
{
    U[0] = P[0];
    U[1] = P[0] * P[1];
    U[2] = P[0] * P[2];
    U[3] = P[0] * P[3];
    U[4] = P[0] * 0.5 * (P[1] * P[1] + P[2] * P[2] + P[3] * P[3]) + P[4] * K.igm1;
}

// P(U) with one reciprocal of the density
__device__ __forceinline__ void prim_of(const Consts& K, const double* U, double* P)
{
    const double ir = rcp(U[0]);
    P[0] = U[0];
    P[1] = U[1] * ir;
    P[2] = U[2] * ir;
    P[3] = U[3] * ir;
    P[4] = (U[4] - 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) * ir) * K.gm1;
}

// CFL time step of a zone: Lmin / max over axes of (|v_axis| + cs), the
// reference's min over zones of L / max |eigenvalue| with axis normals
__device__ __forceinline__ double courant(const Consts& K, const double* P, double Lmin)
{
    const double cs = sound(K, P[0], P[4]);
    const double vm = __builtin_fmax(__builtin_fabs(P[1]), __builtin_fmax(__builtin_fabs(P[2]), __builtin_fabs(P[3])));
    return Lmin * rcp(vm + cs);
}

// Tile of 64 x TY zones marching over z planes; TY = 4 keeps three
// work-groups per CU within the LDS.
constexpr int FTX = 64, FTY = 4;

template <int Riem, bool Z>
static __global__ void __launch_bounds__(FTX * FTY) __attribute__((amdgpu_waves_per_eu(3, 8)))
k_fast_plm(const double* __restrict__ Pin, const double* __restrict__ P0, double* __restrict__ Pout, Geom G,
           double b, int first_stage, int courant_out, Ctl* ctl, int kz)
{
    if (!ctl->active) return;
    constexpr int NG = 2, TX = FTX, TY = FTY;
    constexpr int HX = TX + 2 * NG, HY = TY + 2 * NG;
    constexpr int NT = TX * TY;
    __shared__ double plane[5][HY][HX];   // plane k, halo 2; zone (i, j) at [j - j0 + 2][i - i0 + 2]
    __shared__ double sy[5][TY + 2][TX];  // y half-slopes of rows -1 .. TY at [r + 1]
    __shared__ double sxe[5][TY][3];      // x half-slopes of zones i0 - 1, i0 + 63, i0 + 64
    __shared__ double fy[5][TY + 1][TX];  // y-face fluxes, face j0 + r at [r]
    __shared__ double fxe[5][TY];         // x-face i0 + 64 of each row
    __shared__ double wmin[TX * TY / 64];

    const int ntx = (G.n[0] + TX - 1) / TX, nty = (G.n[1] + TY - 1) / TY;
    const int ntz = Z ? (G.n[2] + kz - 1) / kz : 1;
    const int tile = xcd_tile(blockIdx.x, ntx * nty * ntz);
    const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);

    Consts K;
    K.gamma = G.gamma;
    K.gm1 = G.gamma - 1.0;
    K.igm1 = 1.0 / K.gm1;
    K.th2 = 0.5 * G.theta;
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
    const double A0 = G.dA[0], A1 = G.dA[1], A2 = G.dA[2];
    const double fac = -1.0 / G.dV;
    const double dt = ctl->dt;
    const int ic = min(i, G.N[0] - 1), jc = min(j, G.N[1] - 1);
    const long long col = (long long)jc * G.N[0] + ic;

    double W[5][4];  // z window P(k - 1 .. k + 2) of the zone's column
    double hzc[5];   // z half-slope of zone k
    double Fzlo[5];
    if constexpr (Z)
    {
        double w0[5];
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            w0[q] = Pin[q * G.fs + col + (long long)(kb - 2) * sxy];
#pragma unroll
            for (int m = 0; m < 3; ++m) W[q][m + 1] = Pin[q * G.fs + col + (long long)(kb - 1 + m) * sxy];
        }
        double PL[5], PR[5];
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            const double hlo = half_slope(K.th2, w0[q], W[q][1], W[q][2]);
            hzc[q] = half_slope(K.th2, W[q][1], W[q][2], W[q][3]);
            PL[q] = W[q][1] + hlo;
            PR[q] = W[q][2] - hzc[q];
        }
        face_flux<Riem, 2>(K, PL, PR, Fzlo);
    }

    for (int k = kb; k < ke; ++k)
    {
        const long long pbase = (long long)k * sxy;
        if constexpr (Z)
        {
#pragma unroll
            for (int q = 0; q < 5; ++q) plane[q][row + NG][lane + NG] = W[q][2];
            for (int t = tid; t < HX * HY; t += NT)
            {
                const int hx = t % HX, hy = t / HX;
                if (hx >= NG && hx < NG + TX && hy >= NG && hy < NG + TY) continue;
                const int gi = min(i0 - NG + hx, G.N[0] - 1), gj = min(j0 - NG + hy, G.N[1] - 1);
                const long long c = pbase + (long long)gj * G.N[0] + gi;
#pragma unroll
                for (int q = 0; q < 5; ++q) plane[q][hy][hx] = Pin[q * G.fs + c];
            }
        }
        else
        {
            for (int t = tid; t < HX * HY; t += NT)
            {
                const int hx = t % HX, hy = t / HX;
                const int gi = min(i0 - NG + hx, G.N[0] - 1), gj = min(j0 - NG + hy, G.N[1] - 1);
                const long long c = pbase + (long long)gj * G.N[0] + gi;
#pragma unroll
                for (int q = 0; q < 5; ++q) plane[q][hy][hx] = Pin[q * G.fs + c];
            }
        }
        double Fzhi[5];
        if constexpr (Z)
        {
            double PL[5], PR[5];
#pragma unroll
            for (int q = 0; q < 5; ++q)
            {
                W[q][0] = W[q][1];
                W[q][1] = W[q][2];
                W[q][2] = W[q][3];
                W[q][3] = Pin[q * G.fs + col + (long long)(k + 2) * sxy];
                const double hn = half_slope(K.th2, W[q][1], W[q][2], W[q][3]);
                PL[q] = W[q][1] + hzc[q];
                PR[q] = W[q][2] - hn;
                hzc[q] = hn;
            }
            if (cell_xy) face_flux<Riem, 2>(K, PL, PR, Fzhi);
        }
        __syncthreads();

        // half-slopes: own zone in x (register) and y (LDS); tile-edge zones
        double hx[5];
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

        double Fx[5], Fy[5];
        double hl[5];
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            const double up = __shfl_up(hx[q], 1, 64);
            hl[q] = lane == 0 ? sxe[q][row][0] : up;
        }
        if (xface)
        {
            double PL[5], PR[5];
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
            double PL[5], PR[5];
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
            double PL[5], PR[5], F[5];
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

        double Fxr[5];
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            const double down = __shfl_down(Fx[q], 1, 64);
            Fxr[q] = lane == TX - 1 ? fxe[q][row] : down;
        }
        if (cell_xy)
        {
            const long long c = pbase + col;
            // the stage's own state U(Pin) from the zone's input value (in the z
            // window, or the LDS plane in 2-D); the step's start state U(P0) is
            // the same value in the first stage and read from P0 after it
            double pin[5], Us[5], U0[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) pin[q] = Z ? W[q][1] : plane[q][row + 2][lane + 2];
            cons_of(K, pin, Us);
            if (first_stage)
            {
#pragma unroll
                for (int q = 0; q < 5; ++q) U0[q] = Us[q];
            }
            else
            {
                double p0[5];
#pragma unroll
                for (int q = 0; q < 5; ++q) p0[q] = P0[q * G.fs + c];
                cons_of(K, p0, U0);
            }
            double Un[5];
#pragma unroll
            for (int q = 0; q < 5; ++q)
            {
                double acc = (Fxr[q] - Fx[q]) * A0 + (fy[q][row + 1][lane] - Fy[q]) * A1;
                if constexpr (Z) acc += (Fzhi[q] - Fzlo[q]) * A2;
                const double L = fac * acc;
                Un[q] = U0[q] * (1 - b) + (Us[q] + dt * L) * b;
            }
            double Pn[5];
            prim_of(K, Un, Pn);
#pragma unroll
            for (int q = 0; q < 5; ++q) Pout[q * G.fs + c] = Pn[q];
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
        __syncthreads();
    }
    if (courant_out) courant_to_bins(dmin, wmin, TX * TY / 64, tid, ctl);
}

// the blocks this kernel serves: hydro, PLM, 2-D or 3-D, two or three stages
inline bool supported(const mara_desc* d)
{
    return d->law == MARA_LAW_HYDRO && d->recon == MARA_RECON_PLM && d->cells[1] > 1 && d->rk_order >= 2;
}

} // namespace fast64
} // namespace mara
