// thermal_convection.hip -- a source-term code object for a spherical mesh:
// the source terms of the ThermalConvection program
// (src/Problems/ThermalConvection.cpp:218-245).  The zone's centre arrives
// as (r, theta, phi); the momentum equations of spherical coordinates carry
// the geometric terms, gravity g0 / r^2 (g0 = params[0]) acts on the radial
// momentum and the energy, and the energy gains params[1] exp(-r).  cot is
// tan(pi/2 - theta), as the program writes it.
#include "../../include/mara_source.h"

__device__ inline void thermal_convection(double r, double q, double p, const double* t, const double* P, double* S,
                                          const double* params)
{
    (void)p;
    (void)t;
    const double g0 = params[0], q0 = params[1];
    const double rho = P[0], vr = P[1], vq = P[2], vp = P[3], pre = P[4];
    const double cotq = tan(M_PI_2 - q);
    S[0] = 0.0;
    S[1] = (2 * pre + rho * (vq * vq + vp * vp)) / r;
    S[2] = (pre * cotq + rho * (vp * vp * cotq - vr * vq)) / r;
    S[3] = -rho * vp * (vr + vq * cotq) / r;
    S[4] = 0.0;
    const double g = g0 * pow(r, -2.0);
    S[1] += -rho * g;
    S[4] += -rho * g * vr;
    S[4] += q0 * exp(-r);
}

MARA_SOURCE_KERNEL(mara_source_thermal_convection, thermal_convection)
