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
    fast64::launch<double>(*s.G, s.stream, s.d->riemann, s.Pin, s.P0, s.Pout, s.b, s.first, s.cout, s.ctl, s.kz);
}

void stage_kernels_fast32(const StageArgs& s)
{
    fast64::launch<float>(*s.G, s.stream, s.d->riemann, s.Pin, s.P0, s.Pout, s.b, s.first, s.cout, s.ctl, s.kz);
}

} // namespace mara
