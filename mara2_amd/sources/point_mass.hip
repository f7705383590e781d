// point_mass.hip -- a source-term code object for mara_block_set_source: the
// softened gravity of a point mass GM = params[0], softening length squared
// params[1], at the position (t[0], t[1], t[2]) held in the auxiliary data
// (where the reference passes its star-particle positions), acting on the
// momentum (rho a) and the energy (rho v . a).
#include "../../include/mara_source.h"

__device__ inline void point_mass(double x, double y, double z, const double* t, const double* P, double* S,
                                  const double* params)
{
    const double dx = x - t[0], dy = y - t[1], dz = z - t[2];
    const double r2 = dx * dx + dy * dy + dz * dz + params[1];
    const double f = -params[0] / (r2 * sqrt(r2));
    const double a0 = f * dx, a1 = f * dy, a2 = f * dz;
    S[0] = 0.0;
    S[1] = P[0] * a0;
    S[2] = P[0] * a1;
    S[3] = P[0] * a2;
    S[4] = P[0] * (P[1] * a0 + P[2] * a1 + P[3] * a2);
}

MARA_SOURCE_KERNEL(mara_source_point_mass, point_mass)
