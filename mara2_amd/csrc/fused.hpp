// fused.hpp -- fused stage kernel for the Newtonian hydro law on 2-D and 3-D
// blocks: reconstruction, Riemann fluxes, flux divergence, RK combine and
// primitive recovery in one pass over the zones.
//
// Work decomposition (2.5-D marching):
//   - a work-group owns a tile of TX = 64 (one wavefront per row) by TY = 8
//     interior zones in the x-y plane and marches along z over KZ planes;
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
constexpr int FUSED_KZ = 16;

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

template <int Recon, int Riem, typename R, bool Z>
__global__ void __launch_bounds__(FUSED_TX * FUSED_TY)
k_fused_hydro(const double* __restrict__ Pin, const double* __restrict__ P0, double* __restrict__ U,
              double* __restrict__ Pout, Geom G, double b, int first_stage, int write_u, Ctl* ctl)
{
    constexpr int NG = Recon == 0 ? 1 : Recon == 1 ? 2 : 3;
    constexpr int TX = FUSED_TX, TY = FUSED_TY;
    constexpr int HX = TX + 2 * NG, HY = TY + 2 * NG;
    __shared__ R plane[5][HY][HX];
    __shared__ R fy[5][TY + 1][TX];
    __shared__ R fxe[5][TY];

    if (!ctl->active) return;
    const int lane = threadIdx.x;
    const int row = threadIdx.y;
    const int tid = row * TX + lane;
    const int i0 = G.g[0] + blockIdx.x * TX;
    const int j0 = G.g[1] + blockIdx.y * TY;
    const int i = i0 + lane, j = j0 + row;
    const int ie = G.N[0] - G.g[0], je = G.N[1] - G.g[1];  // one past the last interior zone
    const bool cell_xy = i < ie && j < je;
    const bool xface = i <= ie && j < je;                  // left x-face of (i, j) exists
    const bool yface = j <= je && i < ie;                  // lower y-face of (i, j) exists
    const int kb = Z ? G.g[2] + blockIdx.z * FUSED_KZ : 0;
    const int ke = Z ? min(kb + FUSED_KZ, G.N[2] - G.g[2]) : 1;
    const long long sxy = (long long)G.N[0] * G.N[1];
    const R theta = R(G.theta), gamma = R(G.gamma);
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
            prim_to_cons<0>(G.gamma, p0, U0);
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
            cons_to_prim<0>(G.gamma, G.pfloor, Un, Pn, floor);
#pragma unroll
            for (int q = 0; q < 5; ++q) Pout[q * G.fs + c] = Pn[q];
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
        __syncthreads();
    }
}

// The fused path covers the hydro law on 2-D and 3-D blocks; 1-D blocks and
// the MHD laws (which need the flux arrays for flux-CT) use the flux-array path.
inline bool fused_supported(const mara_desc* d)
{
    return d->law == MARA_LAW_HYDRO && d->cells[1] > 1;
}

inline const char* fused_kernel_name(const mara_desc& d)
{
    return d.precision == MARA_PREC_MIXED ? "k_fused_hydro<float>" : "k_fused_hydro<double>";
}

template <int Recon, int Riem, typename R>
inline void launch_fused_t(const Geom& G, hipStream_t s, const double* Pin, const double* P0, double* U,
                           double* Pout, double b, int first, int write_u, Ctl* ctl)
{
    const bool z = G.N[2] > 1;
    dim3 block(FUSED_TX, FUSED_TY, 1);
    dim3 grid((G.n[0] + FUSED_TX - 1) / FUSED_TX, (G.n[1] + FUSED_TY - 1) / FUSED_TY,
              z ? (G.n[2] + FUSED_KZ - 1) / FUSED_KZ : 1);
    if (z) k_fused_hydro<Recon, Riem, R, true><<<grid, block, 0, s>>>(Pin, P0, U, Pout, G, b, first, write_u, ctl);
    else k_fused_hydro<Recon, Riem, R, false><<<grid, block, 0, s>>>(Pin, P0, U, Pout, G, b, first, write_u, ctl);
}

template <typename R>
inline void launch_fused_r(const mara_desc& d, const Geom& G, hipStream_t s, const double* Pin, const double* P0,
                           double* U, double* Pout, double b, int first, int write_u, Ctl* ctl)
{
    const bool hllc = d.riemann == MARA_RIEMANN_HLLC;
    switch (d.recon)
    {
        case MARA_RECON_PCM:
            if (hllc) launch_fused_t<0, 1, R>(G, s, Pin, P0, U, Pout, b, first, write_u, ctl);
            else launch_fused_t<0, 0, R>(G, s, Pin, P0, U, Pout, b, first, write_u, ctl);
            break;
        case MARA_RECON_PLM:
            if (hllc) launch_fused_t<1, 1, R>(G, s, Pin, P0, U, Pout, b, first, write_u, ctl);
            else launch_fused_t<1, 0, R>(G, s, Pin, P0, U, Pout, b, first, write_u, ctl);
            break;
        default:
            if (hllc) launch_fused_t<2, 1, R>(G, s, Pin, P0, U, Pout, b, first, write_u, ctl);
            else launch_fused_t<2, 0, R>(G, s, Pin, P0, U, Pout, b, first, write_u, ctl);
            break;
    }
}

inline void launch_fused(const mara_desc& d, const Geom& G, hipStream_t s, const double* Pin, const double* P0,
                         double* U, double* Pout, double b, int first, int write_u, Ctl* ctl)
{
    if (d.precision == MARA_PREC_MIXED) launch_fused_r<float>(d, G, s, Pin, P0, U, Pout, b, first, write_u, ctl);
    else launch_fused_r<double>(d, G, s, Pin, P0, U, Pout, b, first, write_u, ctl);
}

} // namespace mara
