// fast.hip -- the fast binary64 form's stage kernel (fused_fast.hpp) and its
// launch.  Compiled with floating-point contraction within expressions
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
    const Geom& G = *s.G;
    const bool z = G.N[2] > 1;
    const bool hllc = s.d->riemann == MARA_RIEMANN_HLLC;
    constexpr int TX = fast64::FTX, TY = fast64::FTY;
    const dim3 block(TX, TY, 1);
    const dim3 grid(((G.n[0] + TX - 1) / TX) * ((G.n[1] + TY - 1) / TY) * (z ? (G.n[2] + s.kz - 1) / s.kz : 1));
    if (z)
    {
        if (hllc)
            fast64::k_fast_plm<1, true><<<grid, block, 0, s.stream>>>(s.Pin, s.P0, s.Pout, G, s.b, s.first, s.cout,
                                                                     s.ctl, s.kz);
        else
            fast64::k_fast_plm<0, true><<<grid, block, 0, s.stream>>>(s.Pin, s.P0, s.Pout, G, s.b, s.first, s.cout,
                                                                     s.ctl, s.kz);
    }
    else
    {
        if (hllc)
            fast64::k_fast_plm<1, false><<<grid, block, 0, s.stream>>>(s.Pin, s.P0, s.Pout, G, s.b, s.first, s.cout,
                                                                      s.ctl, s.kz);
        else
            fast64::k_fast_plm<0, false><<<grid, block, 0, s.stream>>>(s.Pin, s.P0, s.Pout, G, s.b, s.first, s.cout,
                                                                      s.ctl, s.kz);
    }
}

} // namespace mara
