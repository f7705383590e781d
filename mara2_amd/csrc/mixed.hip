// mixed.hip -- the binary32 (mixed-form) instantiations of the stage launches
// of a single block (stage.hpp) and of a batch (batch_state.hpp).  This file
// is compiled with floating-point contraction within expressions (fused
// multiply-add, -ffp-contract=on), which the mixed form allows: its fluxes
// are binary32 and its RK accumulation,
// primitive recovery and flux-CT stay binary64, with no claim to the
// reference's rounding.  The binary64 kernels live in api.hip and batch.hip,
// compiled without contraction.
#include "stage.hpp"
#include "batch_state.hpp"

namespace mara {

void stage_kernels_f32(const StageArgs& s) { stage_kernels<float>(s); }
void stage_round_f32(const StageArgs& s, int* per_cu, int* per_layer) { fused_round_r<float>(*s.d, *s.G, per_cu, per_layer); }

namespace batch {
void run_stage_f32(mara_batch* B, int rk, int lane) { run_stage<float>(B, rk, lane); }
} // namespace batch

} // namespace mara
