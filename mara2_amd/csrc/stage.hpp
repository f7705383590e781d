// stage.hpp -- kernel launches of one RK stage of a single block, templated
// on the flux arithmetic R.  Instantiated for binary64 in api.hip and for
// binary32 in mixed.hip, which is compiled with floating-point contraction
// (the mixed form is not bound to the reference's rounding).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/mara_gpu.h"
#include "kernels.hpp"
#include "fused.hpp"
#include "fused_mhd.hpp"

namespace mara {

// Everything one stage's launches need: the stage reads Pin (neighbours) and
// P0, U (own zone), writes Pout and, when write_u, U.
struct StageArgs
{
    const mara_desc* d;
    const Geom* G;
    hipStream_t stream;
    const double* Pin;
    const double* P0;
    double* U;
    double* Pout;
    void* F[3];
    double b;
    int first, write_u, cout;
    Ctl* ctl;
    bool fused, ct;
    int pencil;  // 1-D pencil tiles of this length, or 0
    bool mhd2d;  // 2-D MHD: fused flux + CT + update tiles
    int kz;      // fused 3-D kernels: z planes per work-group
    int rows;    // flux-array kernels: work-group rows of 64 zones; fused 2-D
                 // MHD tiles: 64 x (4 * rows) zones
    const double* S;  // source terms added to L (same layout as P), or null
    // the fast tiles and the bit-exact PLM tile of a periodic block: read the
    // halos from the interior images of the guard zones, so Pin's guards need
    // not be filled
    bool wrap = false;
    // flux-array path: called between the flux sweeps and the zone update
    // (the viscous-flux hook enqueues its kernel there), or null
    void (*after_flux)(void*);
    // binary64 staging arrays of the fluxes for the hook where the flux
    // arrays are binary32 or flux-CT must come first (null otherwise): the
    // hook then sees and updates these (k_visc_stage), and the zone update
    // reads them (binary64) or their binary32 rounding (k_visc_unstage)
    // without applying flux-CT again
    void* Fv[3] = {nullptr, nullptr, nullptr};
    void* after_flux_ctx;
};

// 2-D work-groups of 64 x rows over an (m0, m1, m2) box: one wavefront per x
// row, z from the grid, so kernels decode indices without integer division.
// rows (1, 2 or 4) is a launch shape; the result does not depend on it.
constexpr int ROWS_DEFAULT = 4;
constexpr int ROWS_SHAPES[3] = {1, 2, 4};
inline dim3 stage_block3(int rows) { return dim3(64, rows, 1); }
inline dim3 stage_grid3(int m0, int m1, int m2, int rows) { return dim3((m0 + 63) / 64, (m1 + rows - 1) / rows, m2); }

template <int Law, int Recon, int Riem, typename R>
void stage_flux_t(const StageArgs& s, int axis)
{
    const Geom& G = *s.G;
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a)
    {
        if (a == axis) { lo[a] = G.g[a]; hi[a] = G.N[a] - G.g[a] + 1; }
        else
        {
            // flux-CT reads the transverse neighbours of each face: one ring more
            const int ring = (s.ct && G.N[a] > 1) ? 1 : 0;
            lo[a] = G.g[a] - ring;
            hi[a] = G.N[a] - G.g[a] + ring;
        }
    }
    const int m0 = hi[0] - lo[0], m1 = hi[1] - lo[1], m2 = hi[2] - lo[2];
    // rows of 64 lanes leave (64 - m0 % 64) lanes of each row's last
    // wavefront idle; past an eighth of the lanes the faces go in one line
    const int padded = (m0 + 63) / 64 * 64;
    const long long faces = (long long)m0 * m1 * m2;
    if (8 * (padded - m0) > padded && faces < (1ll << 31))
        k_flux<Law, Recon, Riem, R><<<(unsigned)((faces + 255) / 256), 256, 0, s.stream>>>(
            s.Pin, (R*)s.F[axis], G, axis, lo[0], lo[1], lo[2], m0, m1, m2, 1, s.ctl);
    else
        k_flux<Law, Recon, Riem, R><<<stage_grid3(m0, m1, m2, s.rows), stage_block3(s.rows), 0, s.stream>>>(
            s.Pin, (R*)s.F[axis], G, axis, lo[0], lo[1], lo[2], m0, m1, m2, 0, s.ctl);
}

template <int Law, typename R>
void stage_flux_law(const StageArgs& s, int axis)
{
    const int recon = s.d->recon, riem = s.d->riemann;
    if (recon == MARA_RECON_PCM)
    {
        if (Law == 0 && riem == 1) stage_flux_t<Law, 0, 1, R>(s, axis);
        else stage_flux_t<Law, 0, 0, R>(s, axis);
    }
    else if (recon == MARA_RECON_PLM)
    {
        if (Law == 0 && riem == 1) stage_flux_t<Law, 1, 1, R>(s, axis);
        else stage_flux_t<Law, 1, 0, R>(s, axis);
    }
    else
    {
        if (Law == 0 && riem == 1) stage_flux_t<Law, 2, 1, R>(s, axis);
        else stage_flux_t<Law, 2, 0, R>(s, axis);
    }
}

template <int Law, typename R>
void stage_update_law(const StageArgs& s)
{
    const Geom& G = *s.G;
    const dim3 grid = stage_grid3(G.n[0], G.n[1], G.n[2], s.rows);
    if (s.ct)
        k_update<Law, R, true><<<grid, stage_block3(s.rows), 0, s.stream>>>(
            s.P0, s.U, s.Pout, (const R*)s.F[0], (const R*)s.F[1], (const R*)s.F[2],
            G, s.b, s.first, s.write_u, s.cout, s.ctl, s.S);
    else
        k_update<Law, R, false><<<grid, stage_block3(s.rows), 0, s.stream>>>(
            s.P0, s.U, s.Pout, (const R*)s.F[0], (const R*)s.F[1], (const R*)s.F[2],
            G, s.b, s.first, s.write_u, s.cout, s.ctl, s.S);
}

template <int Law, int Recon, int Riem, typename R, int PL>
void stage_pencil_pl(const StageArgs& s)
{
    const Geom& G = *s.G;
    const int tiles = (G.n[0] + PL - 2) / (PL - 1);
    if constexpr (Law != 0)
        if (s.ct)
        {
            k_pencil<Law, Recon, Riem, R, true, PL><<<tiles, PL, 0, s.stream>>>(s.Pin, s.P0, s.U, s.Pout, G, s.b,
                                                                               s.first, s.write_u, s.cout, s.ctl, s.S);
            return;
        }
    k_pencil<Law, Recon, Riem, R, false, PL><<<tiles, PL, 0, s.stream>>>(s.Pin, s.P0, s.U, s.Pout, G, s.b, s.first,
                                                                        s.write_u, s.cout, s.ctl, s.S);
}

template <int Law, int Recon, int Riem, typename R>
void stage_pencil_t(const StageArgs& s)
{
    // the relativistic law's pencils (Newton c2p and quartic inlined) come in
    // the default length only
    if constexpr (Law == 2)
    {
        stage_pencil_pl<Law, Recon, Riem, R, PENCIL_DEFAULT>(s);
        return;
    }
    switch (s.pencil)
    {
        case 128: stage_pencil_pl<Law, Recon, Riem, R, 128>(s); break;
        case 512: stage_pencil_pl<Law, Recon, Riem, R, 512>(s); break;
        default: stage_pencil_pl<Law, Recon, Riem, R, 256>(s); break;
    }
}

template <int Law, typename R>
void stage_pencil_law(const StageArgs& s)
{
    const int recon = s.d->recon, riem = s.d->riemann;
    if (recon == MARA_RECON_PCM)
    {
        if (Law == 0 && riem == 1) stage_pencil_t<Law, 0, 1, R>(s);
        else stage_pencil_t<Law, 0, 0, R>(s);
    }
    else if (recon == MARA_RECON_PLM)
    {
        if (Law == 0 && riem == 1) stage_pencil_t<Law, 1, 1, R>(s);
        else stage_pencil_t<Law, 1, 0, R>(s);
    }
    else
    {
        if (Law == 0 && riem == 1) stage_pencil_t<Law, 2, 1, R>(s);
        else stage_pencil_t<Law, 2, 0, R>(s);
    }
}

// The stage's kernels: the fused hydro kernel, the fused 2-D MHD kernel, the
// 1-D pencil kernel, or one flux sweep per axis followed by the zone update
// (all but the fused hydro kernel write the ghost zones themselves).
template <typename R>
void stage_kernels(const StageArgs& s)
{
    if (s.fused)
    {
        launch_fused_r<R>(*s.d, *s.G, s.stream, s.Pin, s.P0, s.U, s.Pout, s.b, s.first, s.write_u, s.cout, s.ctl,
                          s.kz, s.wrap);
        return;
    }
    if (s.mhd2d)
    {
        launch_mhd2d<R>(*s.d, *s.G, s.ct, s.rows, s.stream, s.Pin, s.P0, s.U, s.Pout, s.b, s.first, s.write_u, s.cout, s.ctl,
                        s.S);
        return;
    }
    if (s.pencil)
    {
        switch (s.d->law)
        {
            case 0: stage_pencil_law<0, R>(s); break;
            case 1: stage_pencil_law<1, R>(s); break;
            default: stage_pencil_law<2, R>(s); break;
        }
        return;
    }
    // law 3: NewtonianMHD with cooling (nine fields)
    const int law = s.G->nq == 9 ? 3 : s.d->law;
    for (int axis = 0; axis < 3; ++axis)
    {
        if (s.G->N[axis] < 2) continue;
        switch (law)
        {
            case 0: stage_flux_law<0, R>(s, axis); break;
            case 1: stage_flux_law<1, R>(s, axis); break;
            case 3: stage_flux_law<3, R>(s, axis); break;
            default: stage_flux_law<2, R>(s, axis); break;
        }
    }
    const bool staged = s.after_flux && (s.Fv[0] || s.Fv[1] || s.Fv[2]);
    if (staged)
    {
        for (int axis = 0; axis < 3; ++axis)
            if (s.Fv[axis])
            {
                const long long faces = (long long)(s.G->n[0] + (axis == 0)) * (s.G->n[1] + (axis == 1))
                                        * (s.G->n[2] + (axis == 2));
                k_visc_stage<R><<<(unsigned)((faces + 255) / 256), 256, 0, s.stream>>>(
                    (const R*)s.F[0], (const R*)s.F[1], (const R*)s.F[2], (double*)s.Fv[axis], *s.G, axis, s.ct);
            }
    }
    if (s.after_flux) s.after_flux(s.after_flux_ctx);
    StageArgs u = s;
    if (staged)
    {
        // flux-CT is in the staged fluxes already
        u.ct = false;
        for (int axis = 0; axis < 3; ++axis)
        {
            if (!s.Fv[axis]) continue;
            if constexpr (sizeof(R) == 8) u.F[axis] = s.Fv[axis];
            else
            {
                const long long faces = (long long)(s.G->n[0] + (axis == 0)) * (s.G->n[1] + (axis == 1))
                                        * (s.G->n[2] + (axis == 2));
                k_visc_unstage<R><<<(unsigned)((faces + 255) / 256), 256, 0, s.stream>>>(
                    (const double*)s.Fv[axis], (R*)s.F[axis], *s.G, axis);
            }
        }
    }
    switch (law)
    {
        case 0: stage_update_law<0, R>(u); break;
        case 1: stage_update_law<1, R>(u); break;
        case 3: stage_update_law<3, R>(u); break;
        default: stage_update_law<2, R>(u); break;
    }
}

// binary32 instantiation (mixed.hip)
void stage_kernels_f32(const StageArgs& s);

// the stage kernel of the fast forms (fast.hip) for the blocks it serves
// (fast64_supported: fused hydro + PLM): binary64 faces for the fast
// binary64 form, binary32 faces for the mixed form; the caller runs
// stage_kernels<double> / stage_kernels_f32 for the others
void stage_kernels_fast64(const StageArgs& s);
void stage_kernels_fast32(const StageArgs& s);
// the fused 2-D MHD tiles in the fast binary64 form's arithmetic (fast.hip)
void stage_kernels_fast64_mhd2d(const StageArgs& s);
bool fast64_supported(const mara_desc* d);

// Work-groups resident per CU and tiles per z layer of the fused 3-D kernel a
// stage launches (for the autotuner's z-plane candidates): the fast kernels
// (fast.hip), the binary32 and binary64 fused kernels (mixed.hip, api.hip)
void stage_round_fast(const StageArgs& s, bool f32, int* per_cu, int* per_layer);
void stage_round_f32(const StageArgs& s, int* per_cu, int* per_layer);
void stage_round(const StageArgs& s, bool mixed, int* per_cu, int* per_layer);

// the CFL reduction in a form's own arithmetic (fast.hip; arith one of
// fast64::CflArith other than CFL_EXACT), law 3 for cooled NewtonianMHD
void courant_kernels_form(const double* P, const Geom& G, int law, Ctl* ctl, hipStream_t stream, int arith,
                          unsigned blocks);

} // namespace mara
