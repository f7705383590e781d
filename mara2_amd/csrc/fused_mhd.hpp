// fused_mhd.hpp -- fused stage kernel for the Newtonian MHD law on 2-D blocks
// (z flat): Godunov fluxes, flux-CT, divergence, RK combine, primitive
// recovery and ghost images in one pass, without flux arrays in HBM.
//
// A work-group owns a tile of 64 x TY interior zones (x fastest, one
// wavefront per row).  Phase 1 computes into LDS
//   - the x-face fluxes of the tile's rows and of one row below and above
//     (faces i0 .. i0 + 64), and
//   - the y-face fluxes of the tile's columns and of one column left and
//     right (faces j0 .. j0 + TY),
// each from its stencil read from the stage input (face_flux_at, the
// arithmetic of the flux-array sweep).  The extra row and column are the
// transverse neighbours flux-CT reads (Toth 2000 eq. 25,
// CellCenteredFieldCT::correctGodunovFluxes, src/CellCenteredFieldCT.cpp:14-134).
// Phase 2 forms each zone's divergence from LDS in the order of update_zone
// (hydro fields in the flux arithmetic, B fields through ct_combine in
// binary64) and hands it to finish_zone.  The faces computed are exactly
// those of the flux-array sweeps with their CT ring, so the result is the
// flux-array path's, bit for bit.
#pragma once

#include "kernels.hpp"
#include "fused.hpp"

namespace mara {

// Tile height: binary64 tiles 64 x 4 (LDS 46 KB), binary32 64 x 8 (40 KB).
template <typename R> struct MhdTile
{
    static constexpr int TY = sizeof(R) == 8 ? 4 : 8;
};

template <int Recon, typename R, bool CT>
__device__ __forceinline__ void mhd2d_tile(const double* __restrict__ Pin, const double* __restrict__ P0,
                                           double* __restrict__ U, double* __restrict__ Pout, const Geom& G, double b,
                                           int first_stage, int write_u, int courant_out, Ctl* ctl, int tx, int ty,
                                           const double* __restrict__ Ssrc)
{
    constexpr int NQ = 8;
    constexpr int TX = 64, TY = MhdTile<R>::TY, NT = TX * TY;
    constexpr int FXC = TX + 1;  // x-faces per row
    constexpr int FXR = TY + 2;  // rows j0 - 1 .. j0 + TY
    constexpr int FYC = TX + 2;  // columns i0 - 1 .. i0 + 64
    constexpr int FYR = TY + 1;  // y-faces j0 .. j0 + TY
    __shared__ R fx[NQ][FXR][FXC];
    __shared__ R fy[NQ][FYR][FYC];
    __shared__ double wmin[NT / 64];

    const int lane = threadIdx.x, row = threadIdx.y, tid = row * TX + lane;
    const int i0 = G.g[0] + tx * TX, j0 = G.g[1] + ty * TY;
    const int ie = G.N[0] - G.g[0], je = G.N[1] - G.g[1];  // one past the last interior zone
    const int k = 0;

    // ---- phase 1: face fluxes into LDS ----------------------------------------------
    // x-faces (c, r): face i0 + c of row j0 - 1 + r; with CT the rows j0 - 1 and
    // je feed the B_y correction, without CT only the tile's rows are needed
    const int xr_lo = CT ? 0 : 1;
    const int xr_hi = min(CT ? FXR : FXR - 1, (CT ? je + 1 : je) - (j0 - 1));  // exclusive
    const int xc_hi = min(FXC, ie + 1 - i0);
    for (int t = tid; t < FXR * FXC; t += NT)
    {
        const int c = t % FXC, r = t / FXC;
        if (r < xr_lo || r >= xr_hi || c >= xc_hi) continue;
        R Fq[NQ];
        face_flux_at<1, Recon, 0, R>(Pin, G, 0, i0 + c, j0 - 1 + r, k, Fq, ctl);
#pragma unroll
        for (int q = 0; q < NQ; ++q) fx[q][r][c] = Fq[q];
    }
    // y-faces (c, r): face j0 + r of column i0 - 1 + c
    const int yc_lo = CT ? 0 : 1;
    const int yc_hi = min(CT ? FYC : FYC - 1, (CT ? ie + 1 : ie) - (i0 - 1));
    const int yr_hi = min(FYR, je + 1 - j0);
    for (int t = tid; t < FYR * FYC; t += NT)
    {
        const int c = t % FYC, r = t / FYC;
        if (c < yc_lo || c >= yc_hi || r >= yr_hi) continue;
        R Fq[NQ];
        face_flux_at<1, Recon, 0, R>(Pin, G, 1, i0 - 1 + c, j0 + r, k, Fq, ctl);
#pragma unroll
        for (int q = 0; q < NQ; ++q) fy[q][r][c] = Fq[q];
    }
    __syncthreads();

    // ---- phase 2: zone update ----------------------------------------------------
    const int i = i0 + lane, j = j0 + row;
    double dtc = 1.79769313486231570815e+308;
    if (i < ie && j < je)
    {
        // LDS coordinates: x-faces of the zone at fx[.][row + 1][lane (+1)],
        // y-faces at fy[.][row (+1)][lane + 1]
        const int xr = row + 1, xc = lane, yr = row, yc = lane + 1;
        double L[NQ];
#pragma unroll
        for (int q = 0; q < 5; ++q)
            L[q] = divergence_R<R>(G, fx[q][xr][xc + 1], fx[q][xr][xc], fy[q][yr + 1][yc], fy[q][yr][yc], R(0), R(0));
        if constexpr (CT)
        {
            // B_x (c = 0): y-faces only, Toth with F_y(B_x) along x and F_x(B_y)
            auto cty = [&](int r) {  // y-face j0 + r of the zone's column
                return ct_combine<R>(fy[5][r][yc], fy[5][r][yc + 1], fy[5][r][yc - 1], fx[6][r + 1][xc + 1],
                                     fx[6][r][xc + 1], fx[6][r + 1][xc], fx[6][r][xc]);
            };
            // B_y (c = 1): x-faces only, Toth with F_x(B_y) along y and F_y(B_x)
            auto ctx = [&](int c) {  // x-face i0 + c of the zone's row
                return ct_combine<R>(fx[6][xr][c], fx[6][xr + 1][c], fx[6][xr - 1][c], fy[5][yr + 1][c + 1],
                                     fy[5][yr + 1][c], fy[5][yr][c + 1], fy[5][yr][c]);
            };
            L[5] = divergence_ct(G, 0.0, 0.0, cty(yr + 1), cty(yr), 0.0, 0.0);
            L[6] = divergence_ct(G, ctx(xc + 1), ctx(xc), 0.0, 0.0, 0.0, 0.0);
            // B_z (c = 2): the z axis is flat, so the fluxes stay as they are
            L[7] = divergence_ct(G, double(fx[7][xr][xc + 1]), double(fx[7][xr][xc]), double(fy[7][yr + 1][yc]),
                                 double(fy[7][yr][yc]), 0.0, 0.0);
        }
        else
        {
#pragma unroll
            for (int q = 5; q < NQ; ++q)
                L[q] = divergence_R<R>(G, fx[q][xr][xc + 1], fx[q][xr][xc], fy[q][yr + 1][yc], fy[q][yr][yc], R(0),
                                       R(0));
        }
        dtc = finish_zone<1, R>(P0, U, Pout, G, b, first_stage, write_u, courant_out, ctl, i, j, k, L, Ssrc);
    }
    if (courant_out) courant_to_bins(dtc, wmin, NT / 64, tid, ctl);
}

template <int Recon, typename R, bool CT>
static __global__ void __launch_bounds__(64 * MhdTile<R>::TY)
k_mhd2d(const double* __restrict__ Pin, const double* __restrict__ P0, double* __restrict__ U,
        double* __restrict__ Pout, Geom G, double b, int first_stage, int write_u, int courant_out, Ctl* ctl,
        const double* __restrict__ Ssrc)
{
    if (!ctl->active) return;
    constexpr int TY = MhdTile<R>::TY;
    const int ntx = (G.n[0] + 63) / 64, nty = (G.n[1] + TY - 1) / TY;
    const int tile = xcd_tile(blockIdx.x, ntx * nty);
    mhd2d_tile<Recon, R, CT>(Pin, P0, U, Pout, G, b, first_stage, write_u, courant_out, ctl, tile % ntx, tile / ntx,
                             Ssrc);
}

// 2-D MHD blocks of two or three stages (the z axis flat, x and y fleshed):
// the fused tiles on the fused kernels' buffer schedule (Pin != Pout).
inline bool mhd2d_supported(const mara_desc* d)
{
    return d->law == MARA_LAW_MHD && d->cells[0] > 1 && d->cells[1] > 1 && d->cells[2] == 1 && d->guards[0] > 0
           && d->guards[1] > 0 && d->rk_order >= 2;
}

template <int Recon, typename R>
inline void launch_mhd2d_t(const Geom& G, bool ct, hipStream_t s, const double* Pin, const double* P0, double* U,
                           double* Pout, double b, int first, int write_u, int cout, Ctl* ctl, const double* Ssrc)
{
    constexpr int TY = MhdTile<R>::TY;
    const dim3 block(64, TY, 1);
    const dim3 grid(((G.n[0] + 63) / 64) * ((G.n[1] + TY - 1) / TY));
    if (ct) k_mhd2d<Recon, R, true><<<grid, block, 0, s>>>(Pin, P0, U, Pout, G, b, first, write_u, cout, ctl, Ssrc);
    else k_mhd2d<Recon, R, false><<<grid, block, 0, s>>>(Pin, P0, U, Pout, G, b, first, write_u, cout, ctl, Ssrc);
}

template <typename R>
inline void launch_mhd2d(const mara_desc& d, const Geom& G, bool ct, hipStream_t s, const double* Pin,
                         const double* P0, double* U, double* Pout, double b, int first, int write_u, int cout,
                         Ctl* ctl, const double* Ssrc)
{
    switch (d.recon)
    {
        case MARA_RECON_PCM: launch_mhd2d_t<0, R>(G, ct, s, Pin, P0, U, Pout, b, first, write_u, cout, ctl, Ssrc); break;
        case MARA_RECON_PLM: launch_mhd2d_t<1, R>(G, ct, s, Pin, P0, U, Pout, b, first, write_u, cout, ctl, Ssrc); break;
        default: launch_mhd2d_t<2, R>(G, ct, s, Pin, P0, U, Pout, b, first, write_u, cout, ctl, Ssrc); break;
    }
}

} // namespace mara
