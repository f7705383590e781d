// viscous_constant.hip -- a viscous-flux code object for
// mara_block_set_viscous_flux: constant kinematic viscosity nu = params[0].
// At the face between cells m (below) and p (above) along an axis of cell
// size h, with rho = (rho_m + rho_p) / 2, mu = nu rho, dv = (v_p - v_m) / h
// and vb = (v_m + v_p) / 2 per velocity component:
//   F(momentum c) -= mu dv_c,   F(energy) -= mu (vb . dv).
#include "../../include/mara_source.h"

__device__ inline void viscous_constant(const mara_flux_args& a, int axis, int i, int j, int k)
{
    const long long cp = ((long long)k * a.N[1] + j) * a.N[0] + i;
    const long long st = axis == 0 ? 1 : axis == 1 ? (long long)a.N[0] : (long long)a.N[0] * a.N[1];
    const long long cm = cp - st;
    const long long f = ((long long)k * (a.N[1] + 1) + j) * (a.N[0] + 1) + i;
    const double* P = a.P;
    double* F = a.F[axis];
    const double h = a.dx[axis];
    const double rho = 0.5 * (P[cm] + P[cp]);
    const double mu = a.params[0] * rho;
    double dv[3], vb[3];
    for (int c = 0; c < 3; ++c)
    {
        const double vm = P[(1 + c) * a.fs + cm], vp = P[(1 + c) * a.fs + cp];
        dv[c] = (vp - vm) / h;
        vb[c] = 0.5 * (vm + vp);
    }
    for (int c = 0; c < 3; ++c) F[(1 + c) * a.ffs + f] = F[(1 + c) * a.ffs + f] - mu * dv[c];
    F[4 * a.ffs + f] = F[4 * a.ffs + f] - mu * (vb[0] * dv[0] + vb[1] * dv[1] + vb[2] * dv[2]);
}

MARA_VISCOUS_FLUX_KERNEL(mara_viscous_constant, viscous_constant)
