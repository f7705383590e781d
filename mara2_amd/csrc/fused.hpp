// fused.hpp -- fused stage kernel for the Newtonian hydro law on 2-D and 3-D
// blocks: reconstruction, Riemann fluxes, flux divergence, RK combine and
// primitive recovery in one pass over the zones.
//
// Work decomposition (2.5-D marching):
//   - a work-group owns a tile of TX = 64 (one wavefront per row) by TY
//     interior zones in the x-y plane and marches along z over KZ planes;
//     tiles are numbered so that neighbours run on the same XCD (xcd_tile);
//   - plane k of the stage input, with a halo of NG zones in x and y, is
//     staged in LDS for the x- and y-face stencils;
//   - each zone's own z column (k-NG .. k+NG) lives in registers, so the
//     upper z-face flux is computed from registers and carried to the next
//     plane as the lower one;
//   - every lane computes the flux through the left x-face and the lower
//     y-face of its zone; the right x-face comes from the next lane by a
//     wavefront shuffle, the upper y-face from LDS; the faces on the tile's
//     far edges (one per row in x, one per column in y) are computed in an
//     extra pass.
// Each zone sums its six face terms in the reference's order
// (+Fx1 A - Fx0 A + Fy1 A - Fy0 A + Fz1 A - Fz0 A, src/MeshOperator.cpp:366-372),
// so the result is bit-identical to the flux-array path.
#pragma once

#include "kernels.hpp"
#include "../../include/mara_gpu.h"

namespace mara {

constexpr int FUSED_TX = 64;
constexpr int FUSED_TY = 8;
constexpr int FUSED_KZ = 16;  // default z planes per work-group (a launch shape)
constexpr int KZ_SHAPES[4] = {4, 8, 16, 32};  // the tuner's fixed candidates
constexpr int KZ_MAX = 1024;                   // set_shape accepts 1 .. KZ_MAX

// Work-group -> tile order that keeps neighbouring tiles on one XCD.  The
// dispatcher hands work-group L of a 1-D grid to XCD L % 8, and each XCD has
// its own L2.  Tiles [start(x), start(x + 1)) in x-fastest order go to XCD x,
// in launch order, so tiles that share halo rows and columns run side by side
// behind the same L2.  The map is a bijection on [0, ntiles).
constexpr int NUM_XCD = 8;

__device__ __forceinline__ int xcd_tile(int L, int ntiles)
{
    const int x = L % NUM_XCD, l = L / NUM_XCD;
    const int q = ntiles / NUM_XCD, r = ntiles % NUM_XCD;
    return x * q + (x < r ? x : r) + l;
}

// Whole-wavefront lane shifts by one through DPP (wave_shr:1 / wave_shl:1 on
// every dword): lane l takes v of lane l - 1 (shift_up) or l + 1
// (shift_down), and the lane with no source, 0 or 63, keeps `edge` (bound_ctrl
// off).  One v_mov_b32_dpp per dword, where __shfl_up/__shfl_down take an
// LDS permute and the edge a select.  The fast tiles use these; the
// bit-exact tile keeps the shuffles, against which these measured 0.5 %
// slower there.  Called with all 64 lanes active.
template <int CTRL>
__device__ __forceinline__ float dpp_shift(float v, float edge)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
                                                                 __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double v, double edge)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned long long e = __builtin_bit_cast(unsigned long long, edge);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)e, (int)(unsigned)b, CTRL, 0xf, 0xf, false);
    const unsigned hi =
        (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(e >> 32), (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <typename R> __device__ __forceinline__ R shift_up(R v, R edge) { return dpp_shift<0x138>(v, edge); }
template <typename R> __device__ __forceinline__ R shift_down(R v, R edge) { return dpp_shift<0x130>(v, edge); }

// Index s along axis a of a periodic block, with a guard zone replaced by the
// interior zone whose value the periodic ghost fill copies into it (s < g:
// s + n; s >= g + n: s - n), then clamped to the array (ragged tiles read
// past the mesh into lanes that update nothing).  Tiles that read their
// halos through it do not need the guards filled.
__device__ __forceinline__ int periodic_index(const Geom& G, int a, int s)
{
    const int g = G.g[a], n = G.n[a];
    s = s < g ? s + n : (s >= g + n ? s - n : s);
    return min(max(s, 0), G.N[a] - 1);
}

// Flux through one face from a stencil of 2 NG zones per field (reference
// IntercellFluxScheme + RiemannSolver, src/IntercellFluxSchemes.cpp:45-117).
template <int Recon, int Riem, typename R>
__device__ __forceinline__ void hydro_face_flux(const R (&S)[5][2 * (Recon == 0 ? 1 : Recon == 1 ? 2 : 3)],
                                                R theta, R gamma, const R* n, R* F)
{
    constexpr int NG = Recon == 0 ? 1 : Recon == 1 ? 2 : 3;
    R PL[5], PR[5];
#pragma unroll
    for (int q = 0; q < 5; ++q)
    {
        if constexpr (NG == 1)
        {
            PL[q] = S[q][0];
            PR[q] = S[q][1];
        }
        else if constexpr (NG == 2)
        {
            PL[q] = plm(theta, S[q][0], S[q][1], S[q][2], R(+1.0));
            PR[q] = plm(theta, S[q][1], S[q][2], S[q][3], R(-1.0));
        }
        else
        {
            PL[q] = weno5<R, true>(S[q][0], S[q][1], S[q][2], S[q][3], S[q][4]);
            PR[q] = weno5<R, false>(S[q][1], S[q][2], S[q][3], S[q][4], S[q][5]);
        }
    }
    State<R, 5> SL, SR;
    hydro_state(gamma, PL, n, SL);
    hydro_state(gamma, PR, n, SR);
    if constexpr (Riem == 1) hllc(SL, SR, n, F);
    else hlle<R, 5>(SL, SR, F);
}

// U = U(P0) of the binary32 flux form: p / (gamma - 1) of
// NewtonianHydro::fromPrimitive (src/ConservationLaws.cpp:253-290) as a
// product with 1 / (gamma - 1), which the tile computes once.  The mixed form
// is held to a tolerance, not to the reference's rounding; the binary64 form
// keeps prim_to_cons.
__device__ __forceinline__ void hydro_cons_mixed(double igm1, const double* P, double* U)
{
    const double vv = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
    U[0] = P[0];
    U[1] = P[0] * P[1];
    U[2] = P[0] * P[2];
    U[3] = P[0] * P[3];
    U[4] = P[0] * 0.5 * vv + P[4] * igm1;
}

template <int Recon, int Riem, typename R, bool Z>
static __global__ void __launch_bounds__(FUSED_TX * FUSED_TY)
k_fused_hydro(const double* __restrict__ Pin, const double* __restrict__ P0, double* __restrict__ U,
              double* __restrict__ Pout, Geom G, double b, int first_stage, int write_u, int courant_out, Ctl* ctl,
              int kz)
{
    constexpr int NG = Recon == 0 ? 1 : Recon == 1 ? 2 : 3;
    constexpr int TX = FUSED_TX, TY = FUSED_TY;
    constexpr int HX = TX + 2 * NG, HY = TY + 2 * NG;
    __shared__ R plane[5][HY][HX];
    __shared__ R fy[5][TY + 1][TX];
    __shared__ R fxe[5][TY];
    __shared__ double wmin[TX * TY / 64];

    if (!ctl->active) return;
    double dmin = 1.79769313486231570815e+308;  // this zone column's CFL minimum (last stage)
    const int lane = threadIdx.x;
    const int row = threadIdx.y;
    const int tid = row * TX + lane;
    const int ntx = (G.n[0] + TX - 1) / TX, nty = (G.n[1] + TY - 1) / TY;
    const int ntz = Z ? (G.n[2] + kz - 1) / kz : 1;
    const int tile = xcd_tile(blockIdx.x, ntx * nty * ntz);
    const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
    const int i0 = G.g[0] + tx * TX;
    const int j0 = G.g[1] + ty * TY;
    const int i = i0 + lane, j = j0 + row;
    const int ie = G.N[0] - G.g[0], je = G.N[1] - G.g[1];  // one past the last interior zone
    const bool cell_xy = i < ie && j < je;
    const bool xface = i <= ie && j < je;                  // left x-face of (i, j) exists
    const bool yface = j <= je && i < ie;                  // lower y-face of (i, j) exists
    const int kb = Z ? G.g[2] + tz * kz : 0;
    const int ke = Z ? min(kb + kz, G.N[2] - G.g[2]) : 1;
    const long long sxy = (long long)G.N[0] * G.N[1];
    const R theta = R(G.theta), gamma = R(G.gamma);
    const double igm1 = 1.0 / (G.gamma - 1.0);  // binary32 form's U(P0)
    const R nx[3] = {R(G.nhat[0]), R(G.nhat[1]), R(G.nhat[2])};
    const R ny[3] = {R(G.nhat[3]), R(G.nhat[4]), R(G.nhat[5])};
    const R nz[3] = {R(G.nhat[6]), R(G.nhat[7]), R(G.nhat[8])};
    const R A0 = R(G.dA[0]), A1 = R(G.dA[1]), A2 = R(G.dA[2]);
    const R fac = R(-1.0 / G.dV);
    const double dt = ctl->dt;
    const int ic = min(i, G.N[0] - 1), jc = min(j, G.N[1] - 1);  // clamped for safe loads
    const long long col = (long long)jc * G.N[0] + ic;

    // z window of the zone's own column: W[q][m] = P(k - NG + m), m = 0 .. 2 NG
    R W[5][2 * NG + 1];
    R Fzlo[5];
    if constexpr (Z)
    {
#pragma unroll
        for (int m = 0; m < 2 * NG; ++m)
#pragma unroll
            for (int q = 0; q < 5; ++q) W[q][m + 1] = R(Pin[q * G.fs + col + (long long)(kb - NG + m) * sxy]);
        // lower z-face of the first plane
        R S[5][2 * NG];
#pragma unroll
        for (int q = 0; q < 5; ++q)
#pragma unroll
            for (int m = 0; m < 2 * NG; ++m) S[q][m] = W[q][m + 1];
        hydro_face_flux<Recon, Riem, R>(S, theta, gamma, nz, Fzlo);
    }

    for (int k = kb; k < ke; ++k)
    {
        // ---- stage plane k with its x-y halo in LDS -----------------------------
        const long long pbase = (long long)k * sxy;
        for (int t = tid; t < HX * HY; t += TX * TY)
        {
            const int hx = t % HX, hy = t / HX;
            const int gi = min(i0 - NG + hx, G.N[0] - 1), gj = min(j0 - NG + hy, G.N[1] - 1);
            const long long c = pbase + (long long)gj * G.N[0] + gi;
#pragma unroll
            for (int q = 0; q < 5; ++q) plane[q][hy][hx] = R(Pin[q * G.fs + c]);
        }
        // ---- slide the z window and compute the upper z-face flux -----------------
        R Fzhi[5];
        if constexpr (Z)
        {
#pragma unroll
            for (int q = 0; q < 5; ++q)
            {
#pragma unroll
                for (int m = 0; m < 2 * NG; ++m) W[q][m] = W[q][m + 1];
                W[q][2 * NG] = R(Pin[q * G.fs + col + (long long)(k + NG) * sxy]);
            }
            R S[5][2 * NG];
#pragma unroll
            for (int q = 0; q < 5; ++q)
#pragma unroll
                for (int m = 0; m < 2 * NG; ++m) S[q][m] = W[q][m + 1];
            if (cell_xy) hydro_face_flux<Recon, Riem, R>(S, theta, gamma, nz, Fzhi);
        }
        __syncthreads();

        // ---- left x-face and lower y-face of the zone -------------------------------
        R Fx[5], Fy[5];
        if (xface)
        {
            R S[5][2 * NG];
#pragma unroll
            for (int q = 0; q < 5; ++q)
#pragma unroll
                for (int m = 0; m < 2 * NG; ++m) S[q][m] = plane[q][row + NG][lane + m];
            hydro_face_flux<Recon, Riem, R>(S, theta, gamma, nx, Fx);
        }
        if (yface)
        {
            R S[5][2 * NG];
#pragma unroll
            for (int q = 0; q < 5; ++q)
#pragma unroll
                for (int m = 0; m < 2 * NG; ++m) S[q][m] = plane[q][row + m][lane + NG];
            hydro_face_flux<Recon, Riem, R>(S, theta, gamma, ny, Fy);
#pragma unroll
            for (int q = 0; q < 5; ++q) fy[q][row][lane] = Fy[q];
        }
        // ---- faces on the tile's far edges: x-face TX of each row, y-face TY of each column
        if (tid < TY + TX)
        {
            if (tid < TY)
            {
                const int r = tid;
                if (i0 + TX <= ie && j0 + r < je)
                {
                    R S[5][2 * NG], F[5];
#pragma unroll
                    for (int q = 0; q < 5; ++q)
#pragma unroll
                        for (int m = 0; m < 2 * NG; ++m) S[q][m] = plane[q][r + NG][TX + m];
                    hydro_face_flux<Recon, Riem, R>(S, theta, gamma, nx, F);
#pragma unroll
                    for (int q = 0; q < 5; ++q) fxe[q][r] = F[q];
                }
            }
            else
            {
                const int cx = tid - TY;
                if (j0 + TY <= je && i0 + cx < ie)
                {
                    R S[5][2 * NG], F[5];
#pragma unroll
                    for (int q = 0; q < 5; ++q)
#pragma unroll
                        for (int m = 0; m < 2 * NG; ++m) S[q][m] = plane[q][TY + m][cx + NG];
                    hydro_face_flux<Recon, Riem, R>(S, theta, gamma, ny, F);
#pragma unroll
                    for (int q = 0; q < 5; ++q) fy[q][TY][cx] = F[q];
                }
            }
        }
        __syncthreads();

        // ---- divergence, RK combine, primitive recovery -----------------------------
        R Fxr[5];
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            const R down = __shfl_down(Fx[q], 1, 64);
            Fxr[q] = lane == TX - 1 ? fxe[q][row] : down;
        }
        if (cell_xy)
        {
            const long long c = pbase + col;
            double p0[5], U0[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) p0[q] = P0[q * G.fs + c];
            if constexpr (sizeof(R) == 4) hydro_cons_mixed(igm1, p0, U0);
            else prim_to_cons<0>(G.gamma, p0, U0);
            double Un[5];
#pragma unroll
            for (int q = 0; q < 5; ++q)
            {
                const R fyu = fy[q][row + 1][lane];
                R acc = +Fxr[q] * A0 - Fx[q] * A0 + fyu * A1 - Fy[q] * A1;
                if constexpr (Z) acc = acc + Fzhi[q] * A2 - Fzlo[q] * A2;
                else acc = acc + R(0) * A2 - R(0) * A2;
                const double L = double(fac * acc);
                const double Uq = first_stage ? U0[q] : U[q * G.fs + c];
                Un[q] = U0[q] * (1 - b) + (Uq + dt * L) * b;
            }
            double Pn[5];
            int floor;
            if constexpr (sizeof(R) == 4) cons_to_prim_mixed<0>(G.gamma, G.pfloor, Un, Pn, floor);
            else cons_to_prim<0>(G.gamma, G.pfloor, Un, Pn, floor);
#pragma unroll
            for (int q = 0; q < 5; ++q) Pout[q * G.fs + c] = Pn[q];
            if (courant_out)
            {
                int err = 0;
                const double d = courant_of<0, R>(Pn, G, err);
                if (d < dmin) dmin = d;
            }
            if (write_u)
            {
#pragma unroll
                for (int q = 0; q < 5; ++q) U[q * G.fs + c] = Un[q];
            }
        }
        if constexpr (Z)
        {
#pragma unroll
            for (int q = 0; q < 5; ++q) Fzlo[q] = Fzhi[q];
        }
        // the update reads registers and fy / fxe only, which the next plane
        // rewrites after another barrier every wave must reach: the next
        // plane's staging need not wait here (a 2-D tile has one plane)
    }
    if (courant_out) courant_to_bins(dmin, wmin, TX * TY / 64, tid, ctl);
}

// The zones of plane k that a TX x TY tile (halo NG = 2) stages from HBM into
// its LDS plane[5][TY + 4][TX + 4]: in 3-D the ring around the tile (its own
// zones come from the z window), else the whole plane; zones past the block
// are clamped to its last zone and feed no face that is used.  load() issues
// every load of the plane before store() waits on one, so staging a plane
// costs one memory latency rather than one per pass of a strided loop.
template <typename R, int TX, int TY, bool Z>
struct PlaneRing
{
    static constexpr int NG = 2, HX = TX + 2 * NG, HY = TY + 2 * NG, NT = TX * TY;
    static constexpr int NH = Z ? HX * HY - NT : HX * HY;
    static constexpr int PH = (NH + NT - 1) / NT;
    R v[PH][5];
    int pos[PH];

    // wrap: read the halo of a periodic block from the interior images of
    // its guard zones (periodic_index) instead of the guards
    __device__ __forceinline__ void load(const double* __restrict__ Pin, const Geom& G, long long pbase, int i0,
                                         int j0, int tid, bool wrap = false)
    {
#pragma unroll
        for (int p = 0; p < PH; ++p)
        {
            const int h = tid + p * NT;
            pos[p] = -1;
            if ((p + 1) * NT <= NH || h < NH)
            {
                int hx, hy;
                if constexpr (Z)
                {
                    // rows 0, 1 and TY + 2, TY + 3 whole; columns 0, 1,
                    // HX - 2, HX - 1 of rows 2 .. TY + 1
                    if (h < 4 * HX)
                    {
                        const int r = h / HX;
                        hy = r < 2 ? r : TY + r;
                        hx = h - r * HX;
                    }
                    else
                    {
                        const int e = h - 4 * HX, c = e & 3;
                        hy = NG + (e >> 2);
                        hx = c < 2 ? c : HX - 4 + c;
                    }
                }
                else
                {
                    hy = h / HX;
                    hx = h - hy * HX;
                }
                int off = wrap ? periodic_index(G, 1, j0 - NG + hy) * G.N[0] + periodic_index(G, 0, i0 - NG + hx)
                               : min(j0 - NG + hy, G.N[1] - 1) * G.N[0] + min(i0 - NG + hx, G.N[0] - 1);
                // an opaque copy: the offset is recomputed per plane rather
                // than its five 64-bit field addresses held across the march
                asm volatile("" : "+v"(off));
                const double* src = Pin + pbase + off;
#pragma unroll
                for (int q = 0; q < 5; ++q) v[p][q] = R(src[q * G.fs]);
                pos[p] = hy * HX + hx;
            }
        }
    }

    __device__ __forceinline__ void store(R (*plane)[HY][HX]) const
    {
#pragma unroll
        for (int p = 0; p < PH; ++p)
            if (pos[p] >= 0)
            {
#pragma unroll
                for (int q = 0; q < 5; ++q) (&plane[q][0][0])[pos[p]] = v[p][q];
            }
    }
};

// ---------------------------------------------------------------------------
// PLM variant: every zone's half-slope h = 0.5 * minmod (one per axis and
// field) is computed once and shared by its two faces, where the reference
// evaluates it once per face (src/IntercellFluxSchemes.cpp:102-107).  The
// face states are v + h (PLM_C2R) and v + (-h) (PLM_C2L); since
// (+1 * 0.5) * m and (-1 * 0.5) * m are exact negations of each other, the
// states -- and so the fluxes -- are the reference's bit for bit.
//   x: h of the zone in a register, the left neighbour's by a lane shuffle;
//   y: h of every row (and the rows above and below the tile) in LDS;
//   z: h of the zone above computed from the register window and carried.
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ R half_slope(R theta, R vl, R v0, R vr)
{
    return R(+1.0) * R(0.5) * minmod(theta, vl, v0, vr);
}

template <int Riem, typename R>
__device__ __forceinline__ void hydro_states_flux(const R* PL, const R* PR, R gamma, const R* n, R* F)
{
    State<R, 5> SL, SR;
    hydro_state(gamma, PL, n, SL);
    hydro_state(gamma, PR, n, SR);
    if constexpr (Riem == 1) hllc(SL, SR, n, F);
    else hlle<R, 5>(SL, SR, F);
}

// Tile height: binary64 state needs twice the LDS and more VGPRs per lane, so
// its tiles are 64 x 4 (three work-groups per CU), binary32 ones 64 x 8.
// WAVES is the occupancy the register allocator is held to: binary64 at 3
// waves per SIMD (168 VGPRs), binary32 at 4 (128 VGPRs, no spills; at 5 it
// spills).
template <typename R> struct PlmTile
{
    static constexpr int TY = sizeof(R) == 8 ? 4 : 8;
    static constexpr int WAVES = sizeof(R) == 8 ? 3 : 4;
};

// One tile (tx, ty, tz) of the fused PLM stage; every thread of the
// work-group calls it (it synchronises internally).
template <int Riem, typename R, bool Z>
__device__ __forceinline__ void fused_plm_tile(const double* __restrict__ Pin, const double* __restrict__ P0,
                                               double* __restrict__ U, double* __restrict__ Pout, const Geom& G,
                                               double b, int first_stage, int write_u, int courant_out, Ctl* ctl,
                                               int tx, int ty, int tz, int kz, bool wrap)
{
    constexpr int NG = 2;
    constexpr int TX = FUSED_TX, TY = PlmTile<R>::TY;
    constexpr int HX = TX + 2 * NG, HY = TY + 2 * NG;
    constexpr int NT = TX * TY;
    __shared__ R plane[5][HY][HX];   // plane k, halo 2; zone (i, j) at [j - j0 + 2][i - i0 + 2]
    __shared__ R sy[5][TY + 2][TX];  // y half-slopes of rows -1 .. TY at [r + 1]
    __shared__ R sxe[5][TY][3];      // x half-slopes of zones i0 - 1, i0 + 63, i0 + 64
    __shared__ R fy[5][TY + 1][TX];  // y-face fluxes, face j0 + r at [r]
    __shared__ R fxe[5][TY];         // x-face i0 + 64 of each row
    __shared__ double wmin[TX * TY / 64];

    double dmin = 1.79769313486231570815e+308;  // this zone column's CFL minimum (last stage)
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
    const R theta = R(G.theta), gamma = R(G.gamma);
    const double igm1 = 1.0 / (G.gamma - 1.0);  // binary32 form's U(P0)
    const R nx[3] = {R(G.nhat[0]), R(G.nhat[1]), R(G.nhat[2])};
    const R ny[3] = {R(G.nhat[3]), R(G.nhat[4]), R(G.nhat[5])};
    const R nz[3] = {R(G.nhat[6]), R(G.nhat[7]), R(G.nhat[8])};
    const R A0 = R(G.dA[0]), A1 = R(G.dA[1]), A2 = R(G.dA[2]);
    const R fac = R(-1.0 / G.dV);
    const double dt = ctl->dt;
    // wrap (periodic blocks): halos, columns past the tile and the z
    // window's guard planes from their interior images (periodic_index)
    const int ic = wrap ? periodic_index(G, 0, i) : min(i, G.N[0] - 1);
    const int jc = wrap ? periodic_index(G, 1, j) : min(j, G.N[1] - 1);
    const long long col = (long long)jc * G.N[0] + ic;
    const auto zplane = [&](int kk) { return (long long)(wrap ? periodic_index(G, 2, kk) : kk) * sxy; };

    // z window W[q][m] = P(k - 1 + m), m = 0..3, and the carried half-slope of zone k
    R W[5][4];
    R hzc[5];
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
        // lower z-face of plane kb: zones kb-1 and kb
        R PL[5], PR[5];
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            const R hlo = half_slope(theta, w0[q], W[q][1], W[q][2]);
            hzc[q] = half_slope(theta, W[q][1], W[q][2], W[q][3]);
            PL[q] = W[q][1] + hlo;
            PR[q] = W[q][2] + (-hzc[q]);
        }
        hydro_states_flux<Riem, R>(PL, PR, gamma, nz, Fzlo);
    }

    for (int k = kb; k < ke; ++k)
    {
        const long long pbase = (long long)k * sxy;
        // the zone's offset in the plane, opaque per plane like the ring's
        int colk = (int)col;
        asm volatile("" : "+v"(colk));
        // plane k: the ring from memory and (3-D) the own zone from the z
        // window, loaded two planes ago; the window's next value is loaded
        // with the ring
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
        // z: slide the window to k-1..k+2, half-slope of zone k+1, upper z-face
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
                const R hn = half_slope(theta, W[q][1], W[q][2], W[q][3]);
                PL[q] = W[q][1] + hzc[q];
                PR[q] = W[q][2] + (-hn);
                hzc[q] = hn;
            }
            if (cell_xy) hydro_states_flux<Riem, R>(PL, PR, gamma, nz, Fzhi);
        }
        __syncthreads();

        // ---- half-slopes: own zone in x (register) and y (LDS); tile-edge zones
        R hx[5];
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            hx[q] = half_slope(theta, plane[q][row + 2][lane + 1], plane[q][row + 2][lane + 2], plane[q][row + 2][lane + 3]);
            sy[q][row + 1][lane] = half_slope(theta, plane[q][row + 1][lane + 2], plane[q][row + 2][lane + 2], plane[q][row + 3][lane + 2]);
        }
        if (lane == TX - 1)
        {
#pragma unroll
            for (int q = 0; q < 5; ++q) sxe[q][row][1] = hx[q];
        }
        if (tid < 2 * TX + 2 * TY)
        {
            if (tid < 2 * TX)  // rows -1 and TY
            {
                const int cx = tid % TX;
                const int r = tid < TX ? -1 : TY;
#pragma unroll
                for (int q = 0; q < 5; ++q)
                    sy[q][r + 1][cx] = half_slope(theta, plane[q][r + 1][cx + 2], plane[q][r + 2][cx + 2], plane[q][r + 3][cx + 2]);
            }
            else  // zones i0 - 1 and i0 + 64 of each row
            {
                const int e = tid - 2 * TX;
                const int r = e % TY;
                const int xs = e < TY ? 1 : TX + 2;  // LDS x index of the zone
#pragma unroll
                for (int q = 0; q < 5; ++q)
                    sxe[q][r][e < TY ? 0 : 2] = half_slope(theta, plane[q][r + 2][xs - 1], plane[q][r + 2][xs], plane[q][r + 2][xs + 1]);
            }
        }
        __syncthreads();

        // ---- left x-face and lower y-face of the zone; tile-edge faces ---------------
        R Fx[5], Fy[5];
        R hl[5];
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            const R up = __shfl_up(hx[q], 1, 64);
            hl[q] = lane == 0 ? sxe[q][row][0] : up;
        }
        if (xface)
        {
            R PL[5], PR[5];
#pragma unroll
            for (int q = 0; q < 5; ++q)
            {
                PL[q] = plane[q][row + 2][lane + 1] + hl[q];
                PR[q] = plane[q][row + 2][lane + 2] + (-hx[q]);
            }
            hydro_states_flux<Riem, R>(PL, PR, gamma, nx, Fx);
        }
        if (yface)
        {
            R PL[5], PR[5];
#pragma unroll
            for (int q = 0; q < 5; ++q)
            {
                PL[q] = plane[q][row + 1][lane + 2] + sy[q][row][lane];
                PR[q] = plane[q][row + 2][lane + 2] + (-sy[q][row + 1][lane]);
            }
            hydro_states_flux<Riem, R>(PL, PR, gamma, ny, Fy);
#pragma unroll
            for (int q = 0; q < 5; ++q) fy[q][row][lane] = Fy[q];
        }
        // the edge faces go to the last waves; the edge slopes went to the first ones
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
                        PR[q] = plane[q][r + 2][TX + 2] + (-sxe[q][r][2]);
                    }
                    hydro_states_flux<Riem, R>(PL, PR, gamma, nx, F);
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
                        PR[q] = plane[q][TY + 2][cx + 2] + (-sy[q][TY + 1][cx]);
                    }
                    hydro_states_flux<Riem, R>(PL, PR, gamma, ny, F);
#pragma unroll
                    for (int q = 0; q < 5; ++q) fy[q][TY][cx] = F[q];
                }
            }
        }
        __syncthreads();

        // ---- divergence, RK combine, primitive recovery ------------------------------
        R Fxr[5];
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            const R down = __shfl_down(Fx[q], 1, 64);
            Fxr[q] = lane == TX - 1 ? fxe[q][row] : down;
        }
        if (cell_xy)
        {
            const long long c = pbase + colk;
            double p0[5], U0[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) p0[q] = P0[q * G.fs + c];
            if constexpr (sizeof(R) == 4) hydro_cons_mixed(igm1, p0, U0);
            else prim_to_cons<0>(G.gamma, p0, U0);
            double Un[5];
#pragma unroll
            for (int q = 0; q < 5; ++q)
            {
                const R fyu = fy[q][row + 1][lane];
                R acc = +Fxr[q] * A0 - Fx[q] * A0 + fyu * A1 - Fy[q] * A1;
                if constexpr (Z) acc = acc + Fzhi[q] * A2 - Fzlo[q] * A2;
                else acc = acc + R(0) * A2 - R(0) * A2;
                const double L = double(fac * acc);
                const double Uq = first_stage ? U0[q] : U[q * G.fs + c];
                Un[q] = U0[q] * (1 - b) + (Uq + dt * L) * b;
            }
            double Pn[5];
            int floor;
            if constexpr (sizeof(R) == 4) cons_to_prim_mixed<0>(G.gamma, G.pfloor, Un, Pn, floor);
            else cons_to_prim<0>(G.gamma, G.pfloor, Un, Pn, floor);
#pragma unroll
            for (int q = 0; q < 5; ++q) Pout[q * G.fs + c] = Pn[q];
            if (courant_out)
            {
                int err = 0;
                const double d = courant_of<0, R>(Pn, G, err);
                if (d < dmin) dmin = d;
            }
            if (write_u)
            {
#pragma unroll
                for (int q = 0; q < 5; ++q) U[q * G.fs + c] = Un[q];
            }
        }
        if constexpr (Z)
        {
#pragma unroll
            for (int q = 0; q < 5; ++q) Fzlo[q] = Fzhi[q];
        }
        // the update reads registers and fy / fxe only, which the next plane
        // rewrites after another barrier every wave must reach: the next
        // plane's staging need not wait here (a 2-D tile has one plane)
    }
    if (courant_out) courant_to_bins(dmin, wmin, TX * TY / 64, tid, ctl);
}

template <int Riem, typename R, bool Z>
static __global__ void __launch_bounds__(FUSED_TX * PlmTile<R>::TY) __attribute__((amdgpu_waves_per_eu(PlmTile<R>::WAVES, 8)))
k_fused_plm(const double* __restrict__ Pin, const double* __restrict__ P0, double* __restrict__ U,
            double* __restrict__ Pout, Geom G, double b, int first_stage, int write_u, int courant_out, Ctl* ctl,
            int kz, int wrap)
{
    if (!ctl->active) return;
    constexpr int TX = FUSED_TX, TY = PlmTile<R>::TY;
    const int ntx = (G.n[0] + TX - 1) / TX, nty = (G.n[1] + TY - 1) / TY;
    const int ntz = Z ? (G.n[2] + kz - 1) / kz : 1;
    const int tile = xcd_tile(blockIdx.x, ntx * nty * ntz);
    const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
    fused_plm_tile<Riem, R, Z>(Pin, P0, U, Pout, G, b, first_stage, write_u, courant_out, ctl, tx, ty, tz, kz,
                               wrap != 0);
}

// The fused path covers the hydro law on 2-D and 3-D blocks; 1-D blocks and
// the MHD laws (which need the flux arrays for flux-CT) use the flux-array path.
// The fused kernel reads neighbouring zones of its input while other
// work-groups write their output, so a stage needs a second buffer; with two
// or three stages the step still ends in the state buffer (Pa -> Pb -> Pa,
// Pa -> Pb -> Pc -> Pa).  A one-stage (RK1) step would end in the other
// buffer, so RK1 blocks take the flux-array path, which updates in place.
inline bool fused_supported(const mara_desc* d)
{
    return d->law == MARA_LAW_HYDRO && d->cells[1] > 1 && d->rk_order >= 2 && d->geometry == MARA_GEOM_CARTESIAN;
}

// 1-D blocks of two or three stages run each stage as pencil tiles
// (k_pencil): fluxes and update in one pass, on the fused kernels' buffer
// schedule, so they hold no flux arrays either.
inline bool pencil_supported(const mara_desc* d)
{
    return d->cells[1] == 1 && d->cells[2] == 1 && d->cells[0] > 1 && d->guards[0] > 0 && d->rk_order >= 2
           && !(d->cooling_rate > 0.0) && d->geometry == MARA_GEOM_CARTESIAN;
}

inline const char* fused_kernel_name(const mara_desc& d)
{
    const bool mixed = d.precision == MARA_PREC_MIXED;
    if (d.precision == MARA_PREC_FP64_FAST && d.recon == MARA_RECON_PLM) return "k_fast_plm<double>";
    if (d.recon == MARA_RECON_PLM) return mixed ? "k_fast_plm<float>" : "k_fused_plm<double>";
    return mixed ? "k_fused_hydro<float>" : "k_fused_hydro<double>";
}

// kz: z planes per work-group (1 .. KZ_MAX; a launch shape, the result does
// not depend on it)
template <int Recon, int Riem, typename R>
inline void launch_fused_t(const Geom& G, hipStream_t s, const double* Pin, const double* P0, double* U,
                           double* Pout, double b, int first, int write_u, int cout, Ctl* ctl, int kz, int wrap)
{
    const bool z = G.N[2] > 1;
    dim3 block(FUSED_TX, FUSED_TY, 1);
    dim3 grid(((G.n[0] + FUSED_TX - 1) / FUSED_TX) * ((G.n[1] + FUSED_TY - 1) / FUSED_TY)
              * (z ? (G.n[2] + kz - 1) / kz : 1));
    if constexpr (Recon == 1)
    {
        constexpr int TY = PlmTile<R>::TY;
        const dim3 pblock(FUSED_TX, TY, 1);
        const dim3 pgrid(((G.n[0] + FUSED_TX - 1) / FUSED_TX) * ((G.n[1] + TY - 1) / TY)
                         * (z ? (G.n[2] + kz - 1) / kz : 1));
        // wrap: the PLM tile reads a periodic block's halos from the interior
        if (z)
            k_fused_plm<Riem, R, true>
                <<<pgrid, pblock, 0, s>>>(Pin, P0, U, Pout, G, b, first, write_u, cout, ctl, kz, wrap);
        else
            k_fused_plm<Riem, R, false>
                <<<pgrid, pblock, 0, s>>>(Pin, P0, U, Pout, G, b, first, write_u, cout, ctl, kz, wrap);
    }
    else
    {
        if (z) k_fused_hydro<Recon, Riem, R, true><<<grid, block, 0, s>>>(Pin, P0, U, Pout, G, b, first, write_u, cout, ctl, kz);
        else k_fused_hydro<Recon, Riem, R, false><<<grid, block, 0, s>>>(Pin, P0, U, Pout, G, b, first, write_u, cout, ctl, kz);
    }
}

// The fused 3-D kernel's work-groups resident on one CU (the occupancy its
// registers and LDS allow) and its tiles per z layer of the block, from which
// the autotuner derives z-plane counts whose tile count fills whole rounds of
// the device.
template <int Recon, int Riem, typename R>
inline void fused_round_t(const Geom& G, int* per_cu, int* per_layer)
{
    int n = 0;
    if constexpr (Recon == 1)
    {
        constexpr int TY = PlmTile<R>::TY;
        *per_layer = ((G.n[0] + FUSED_TX - 1) / FUSED_TX) * ((G.n[1] + TY - 1) / TY);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_plm<Riem, R, true>, FUSED_TX * TY, 0) != hipSuccess)
            n = 0;
    }
    else
    {
        *per_layer = ((G.n[0] + FUSED_TX - 1) / FUSED_TX) * ((G.n[1] + FUSED_TY - 1) / FUSED_TY);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_hydro<Recon, Riem, R, true>, FUSED_TX * FUSED_TY, 0)
            != hipSuccess)
            n = 0;
    }
    (void)hipGetLastError();
    *per_cu = n;
}

template <typename R>
inline void fused_round_r(const mara_desc& d, const Geom& G, int* per_cu, int* per_layer)
{
    const bool hllc = d.riemann == MARA_RIEMANN_HLLC;
    switch (d.recon)
    {
        case MARA_RECON_PCM:
            if (hllc) fused_round_t<0, 1, R>(G, per_cu, per_layer);
            else fused_round_t<0, 0, R>(G, per_cu, per_layer);
            break;
        case MARA_RECON_PLM:
            if (hllc) fused_round_t<1, 1, R>(G, per_cu, per_layer);
            else fused_round_t<1, 0, R>(G, per_cu, per_layer);
            break;
        default:
            if (hllc) fused_round_t<2, 1, R>(G, per_cu, per_layer);
            else fused_round_t<2, 0, R>(G, per_cu, per_layer);
            break;
    }
}

template <typename R>
inline void launch_fused_r(const mara_desc& d, const Geom& G, hipStream_t s, const double* Pin, const double* P0,
                           double* U, double* Pout, double b, int first, int write_u, int cout, Ctl* ctl, int kz,
                           int wrap = 0)
{
    const bool hllc = d.riemann == MARA_RIEMANN_HLLC;
    switch (d.recon)
    {
        case MARA_RECON_PCM:
            if (hllc) launch_fused_t<0, 1, R>(G, s, Pin, P0, U, Pout, b, first, write_u, cout, ctl, kz, wrap);
            else launch_fused_t<0, 0, R>(G, s, Pin, P0, U, Pout, b, first, write_u, cout, ctl, kz, wrap);
            break;
        case MARA_RECON_PLM:
            if (hllc) launch_fused_t<1, 1, R>(G, s, Pin, P0, U, Pout, b, first, write_u, cout, ctl, kz, wrap);
            else launch_fused_t<1, 0, R>(G, s, Pin, P0, U, Pout, b, first, write_u, cout, ctl, kz, wrap);
            break;
        default:
            if (hllc) launch_fused_t<2, 1, R>(G, s, Pin, P0, U, Pout, b, first, write_u, cout, ctl, kz, wrap);
            else launch_fused_t<2, 0, R>(G, s, Pin, P0, U, Pout, b, first, write_u, cout, ctl, kz, wrap);
            break;
    }
}

} // namespace mara
