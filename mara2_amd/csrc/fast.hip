// fast.hip -- the stage kernel of the fast forms (fused_fast.hpp) for single
// blocks: binary64 faces for the fast binary64 form, binary32 faces for the
// mixed form.  Compiled with floating-point contraction within expressions
// (-ffp-contract=on): a multiply-add written in one expression becomes an
// FMA, the same in every inlined copy of a face or zone routine, so the
// result does not depend on the launch shape.  Only names of the fast64
// namespace are instantiated here; the plain binary64 kernels stay in
// api.hip and batch.hip, compiled without contraction.
#include "stage.hpp"
#include "fused_fast.hpp"

namespace mara {

bool fast64_supported(const mara_desc* d) { return fast64::supported(d); }

void stage_kernels_fast64(const StageArgs& s)
{
    fast64::launch<double>(*s.G, s.stream, s.d->riemann, s.Pin, s.P0, s.Pout, s.b, s.first, s.cout, s.ctl, s.kz,
                          s.wrap);
}

void stage_kernels_fast32(const StageArgs& s)
{
    fast64::launch<float>(*s.G, s.stream, s.d->riemann, s.Pin, s.P0, s.Pout, s.b, s.first, s.cout, s.ctl, s.kz,
                          s.wrap);
}

// 2-D MHD blocks of the fast binary64 form: the fused MHD + CT tiles with the
// relativistic law's arithmetic (refined reciprocals and square roots,
// min/max extremes and minmod, one reciprocal of the density in the
// recovery) and contraction
void stage_kernels_fast64_mhd2d(const StageArgs& s)
{
    launch_mhd2d<double, Rel>(*s.d, *s.G, s.ct, s.rows, s.stream, s.Pin, s.P0, s.U, s.Pout, s.b, s.first, s.write_u,
                              s.cout, s.ctl, s.S);
}

void stage_round_fast(const StageArgs& s, bool f32, int* per_cu, int* per_layer)
{
    if (f32) fast64::tile_round<float>(*s.G, s.d->riemann, per_cu, per_layer);
    else fast64::tile_round<double>(*s.G, s.d->riemann, per_cu, per_layer);
}

// The CFL reduction in the arithmetic of a form other than the reference's
// (fast64::CflArith), compiled with the same flags as the stage kernels
// that produce the form's later time steps.
void courant_kernels_form(const double* P, const Geom& G, int law, Ctl* ctl, hipStream_t stream, int arith,
                          unsigned blocks)
{
    switch (law)
    {
        case 0: fast64::k_courant_form<0><<<blocks, 256, 0, stream>>>(P, G, ctl, arith); break;
        case 1: fast64::k_courant_form<1><<<blocks, 256, 0, stream>>>(P, G, ctl, arith); break;
        case 3: fast64::k_courant_form<3><<<blocks, 256, 0, stream>>>(P, G, ctl, arith); break;
        default: fast64::k_courant_form<2><<<blocks, 256, 0, stream>>>(P, G, ctl, arith); break;
    }
}

} // namespace mara
