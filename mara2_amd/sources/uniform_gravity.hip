// uniform_gravity.hip -- a source-term code object for mara_block_set_source:
// a uniform gravitational acceleration g = params[0..2] acting on the
// momentum (rho g) and the energy (rho v . g) of the hydro and MHD laws.
#include "../../include/mara_source.h"

__device__ inline void uniform_gravity(double, double, double, const double*, const double* P, double* S,
                                       const double* params)
{
    const double g0 = params[0], g1 = params[1], g2 = params[2];
    S[0] = 0.0;
    S[1] = P[0] * g0;
    S[2] = P[0] * g1;
    S[3] = P[0] * g2;
    S[4] = P[0] * (P[1] * g0 + P[2] * g1 + P[3] * g2);
}

MARA_SOURCE_KERNEL(mara_source_uniform_gravity, uniform_gravity)
