// fused_mhd.hpp -- fused stage kernel for the Newtonian MHD law on 2-D blocks
// (z flat): Godunov fluxes, flux-CT, divergence, RK combine, primitive
// recovery and ghost images in one pass, without flux arrays in HBM.
//
// A work-group owns a tile of 64 x TY interior zones (x fastest, one
// wavefront per row; TY = 4, 8 or 16, a launch shape).  Phase 1 computes
// into LDS
//   - the x-face fluxes of the tile's rows and of one row below and above
//     (faces i0 .. i0 + 64), and
//   - the y-face fluxes of the tile's columns and of one column left and
//     right (faces j0 .. j0 + TY),
// each from its stencil read from the stage input (face_flux_at, the
// arithmetic of the flux-array sweep), as one list of faces dealt out to
// the work-group's lanes.  The extra row and column are the
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

// Tile height TY = 4 * rows for the launch shape's rows (1, 2 or 4).  LDS per
// work-group (PLM): binary64 46 / 80 / 147 KB; binary32, with its staged
// stencil zones, 40 / 66 / 116 KB.  64 x 8
// tiles are the default in both forms: at 2048^2 their stage takes 0.51 ms
// (binary64) and 0.33 ms (binary32) against 0.60 / 0.35 ms for 64 x 4 and
// 0.54 / 0.36 ms for 64 x 16 (one work-group per CU).
constexpr int MHD_ROWS_DEFAULT = 2;
inline constexpr int mhd_tile_height(int rows) { return 4 * rows; }

// A: the arithmetic of the face states and the zone recovery (Math<R>; Rel
// in the fast binary64 form)
template <int Recon, typename R, bool CT, int TY, typename A = Math<R>>
__device__ __forceinline__ void mhd2d_tile(const double* __restrict__ Pin, const double* __restrict__ P0,
                                           double* __restrict__ U, double* __restrict__ Pout, const Geom& G, double b,
                                           int first_stage, int write_u, int courant_out, Ctl* ctl, int tx, int ty,
                                           const double* __restrict__ Ssrc)
{
    constexpr int NQ = 8;
    constexpr int TX = 64, NT = TX * TY;
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

    // ---- phase 0 (binary32): the tile's stencil zones into LDS ----------------------
    // zones i0 - NG .. i0 + 63 + NG of rows j0 - NG .. j0 + TY - 1 + NG, converted
    // once; indices past the block are clamped and feed no face that is used
    constexpr int NG = Stencil<Recon>::NG;
    constexpr bool STAGE = sizeof(R) == 4;
    constexpr int HX = TX + 2 * NG, HY = TY + 2 * NG;
    __shared__ R sP[STAGE ? NQ : 1][STAGE ? HY : 1][STAGE ? HX : 1];
    if constexpr (STAGE)
    {
        // every load issued before the first LDS write waits on one
        constexpr int PH = (HX * HY + NT - 1) / NT;
        R hv[PH][NQ];
#pragma unroll
        for (int p = 0; p < PH; ++p)
        {
            const int t = tid + p * NT;
            if ((p + 1) * NT <= HX * HY || t < HX * HY)
            {
                const int hx = t % HX, hy = t / HX;
                const int gi = min(i0 - NG + hx, G.N[0] - 1), gj = min(j0 - NG + hy, G.N[1] - 1);
                const long long c = cell_index(G, gi, gj, k);
#pragma unroll
                for (int q = 0; q < NQ; ++q) hv[p][q] = R(Pin[q * G.fs + c]);
            }
        }
#pragma unroll
        for (int p = 0; p < PH; ++p)
        {
            const int t = tid + p * NT;
            if ((p + 1) * NT <= HX * HY || t < HX * HY)
            {
#pragma unroll
                for (int q = 0; q < NQ; ++q) (&sP[q][0][0])[t] = hv[p][q];
            }
        }
        __syncthreads();
    }

    // ---- phase 1: face fluxes into LDS ----------------------------------------------
    // x-faces (c, r): face i0 + c of row j0 - 1 + r; with CT the rows j0 - 1 and
    // je feed the B_y correction, without CT only the tile's rows are needed
    const int xr_lo = CT ? 0 : 1;
    const int xr_hi = min(CT ? FXR : FXR - 1, (CT ? je + 1 : je) - (j0 - 1));  // exclusive
    const int xc_hi = min(FXC, ie + 1 - i0);
    // y-faces (c, r): face j0 + r of column i0 - 1 + c
    const int yc_lo = CT ? 0 : 1;
    const int yc_hi = min(CT ? FYC : FYC - 1, (CT ? ie + 1 : ie) - (i0 - 1));
    const int yr_hi = min(FYR, je + 1 - j0);
    // the faces needed, x-faces first, row by row, dealt out to the lanes as
    // one list; every extent is >= 2 since the tile holds an interior zone.
    // Each kind keeps its own call, so its stencil stride stays a constant.
    const int nxw = xc_hi, nx = (xr_hi - xr_lo) * nxw;
    const int nyw = yc_hi - yc_lo, nf = nx + yr_hi * nyw;
    for (int t = tid; t < nf; t += NT)
    {
        const bool xf = t < nx;
        const int u = xf ? t : t - nx, w = xf ? nxw : nyw;
        const int r = u / w, c = u - r * w;
        R Fq[NQ];
        if (xf)
        {
            const int jj = j0 - 1 + xr_lo + r;
            if constexpr (STAGE)
            {
                // stencil zones i0 + c - NG .. i0 + c + NG - 1 of row jj
                R S[NQ][2 * NG];
                const int ly = jj - (j0 - NG);
#pragma unroll
                for (int q = 0; q < NQ; ++q)
#pragma unroll
                    for (int m = 0; m < 2 * NG; ++m) S[q][m] = sP[q][ly][c + m];
                face_flux<1, Recon, 0, R, A>(S, G, 0, Fq, ctl);
            }
            else
                face_flux_at<1, Recon, 0, R, A>(Pin, G, 0, i0 + c, jj, k, Fq, ctl);
#pragma unroll
            for (int q = 0; q < NQ; ++q) fx[q][xr_lo + r][c] = Fq[q];
        }
        else
        {
            const int ii = i0 - 1 + yc_lo + c;
            if constexpr (STAGE)
            {
                // stencil rows j0 + r - NG .. j0 + r + NG - 1 of column ii
                R S[NQ][2 * NG];
                const int lx = ii - (i0 - NG);
#pragma unroll
                for (int q = 0; q < NQ; ++q)
#pragma unroll
                    for (int m = 0; m < 2 * NG; ++m) S[q][m] = sP[q][r + m][lx];
                face_flux<1, Recon, 0, R, A>(S, G, 1, Fq, ctl);
            }
            else
                face_flux_at<1, Recon, 0, R, A>(Pin, G, 1, ii, j0 + r, k, Fq, ctl);
#pragma unroll
            for (int q = 0; q < NQ; ++q) fy[q][r][yc_lo + c] = Fq[q];
        }
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
        dtc = finish_zone<1, R, A>(P0, U, Pout, G, b, first_stage, write_u, courant_out, ctl, i, j, k, L, Ssrc);
    }
    if (courant_out) courant_to_bins(dtc, wmin, NT / 64, tid, ctl);
}

template <int Recon, typename R, bool CT, int TY, typename A = Math<R>>
static __global__ void __launch_bounds__(64 * TY)
k_mhd2d(const double* __restrict__ Pin, const double* __restrict__ P0, double* __restrict__ U,
        double* __restrict__ Pout, Geom G, double b, int first_stage, int write_u, int courant_out, Ctl* ctl,
        const double* __restrict__ Ssrc)
{
    if (!ctl->active) return;
    const int ntx = (G.n[0] + 63) / 64, nty = (G.n[1] + TY - 1) / TY;
    const int tile = xcd_tile(blockIdx.x, ntx * nty);
    mhd2d_tile<Recon, R, CT, TY, A>(Pin, P0, U, Pout, G, b, first_stage, write_u, courant_out, ctl, tile % ntx,
                                    tile / ntx, Ssrc);
}

// 2-D MHD blocks of two or three stages (the z axis flat, x and y fleshed):
// the fused tiles on the fused kernels' buffer schedule (Pin != Pout).
inline bool mhd2d_supported(const mara_desc* d)
{
    return d->law == MARA_LAW_MHD && d->cells[0] > 1 && d->cells[1] > 1 && d->cells[2] == 1 && d->guards[0] > 0
           && d->guards[1] > 0 && d->rk_order >= 2 && !(d->cooling_rate > 0.0) && d->geometry == MARA_GEOM_CARTESIAN;
}

template <int Recon, typename R, int TY, typename A = Math<R>>
inline void launch_mhd2d_ty(const Geom& G, bool ct, hipStream_t s, const double* Pin, const double* P0, double* U,
                            double* Pout, double b, int first, int write_u, int cout, Ctl* ctl, const double* Ssrc)
{
    const dim3 block(64, TY, 1);
    const dim3 grid(((G.n[0] + 63) / 64) * ((G.n[1] + TY - 1) / TY));
    if (ct)
        k_mhd2d<Recon, R, true, TY, A><<<grid, block, 0, s>>>(Pin, P0, U, Pout, G, b, first, write_u, cout, ctl, Ssrc);
    else
        k_mhd2d<Recon, R, false, TY, A><<<grid, block, 0, s>>>(Pin, P0, U, Pout, G, b, first, write_u, cout, ctl, Ssrc);
}

template <int Recon, typename R, typename A = Math<R>>
inline void launch_mhd2d_t(const Geom& G, bool ct, int rows, hipStream_t s, const double* Pin, const double* P0,
                           double* U, double* Pout, double b, int first, int write_u, int cout, Ctl* ctl,
                           const double* Ssrc)
{
    switch (rows)
    {
        case 1: launch_mhd2d_ty<Recon, R, 4, A>(G, ct, s, Pin, P0, U, Pout, b, first, write_u, cout, ctl, Ssrc); break;
        case 2: launch_mhd2d_ty<Recon, R, 8, A>(G, ct, s, Pin, P0, U, Pout, b, first, write_u, cout, ctl, Ssrc); break;
        default: launch_mhd2d_ty<Recon, R, 16, A>(G, ct, s, Pin, P0, U, Pout, b, first, write_u, cout, ctl, Ssrc); break;
    }
}

template <typename R, typename A = Math<R>>
inline void launch_mhd2d(const mara_desc& d, const Geom& G, bool ct, int rows, hipStream_t s, const double* Pin,
                         const double* P0, double* U, double* Pout, double b, int first, int write_u, int cout,
                         Ctl* ctl, const double* Ssrc)
{
    switch (d.recon)
    {
        case MARA_RECON_PCM:
            launch_mhd2d_t<0, R, A>(G, ct, rows, s, Pin, P0, U, Pout, b, first, write_u, cout, ctl, Ssrc);
            break;
        case MARA_RECON_PLM:
            launch_mhd2d_t<1, R, A>(G, ct, rows, s, Pin, P0, U, Pout, b, first, write_u, cout, ctl, Ssrc);
            break;
        default: launch_mhd2d_t<2, R, A>(G, ct, rows, s, Pin, P0, U, Pout, b, first, write_u, cout, ctl, Ssrc); break;
    }
}

} // namespace mara
