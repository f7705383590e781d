// ref_bridge.cpp -- extern "C" entry points over the reference's own
// standalone translation units (Reconstruction, QuarticPolynomial,
// ConsToPrimRMHD, UnitVector), so tests can compare the oracle with the
// reference bit-for-bit.  TEST INFRASTRUCTURE ONLY: built by oracle/Makefile
// into oracle/_ref/ next to the reference sources it is compiled with; the
// reference sources themselves are compiled where they lie and never copied.
#include "Reconstruction.hpp"
#include "QuarticPolynomial.hpp"
#include "ConsToPrimRMHD.hpp"
#include "UnitVector.hpp"

extern "C" {

void ref_unit_vectors(double out[9])
{
    const UnitVector* v[3] = {&UnitVector::xhat, &UnitVector::yhat, &UnitVector::zhat};
    for (int a = 0; a < 3; ++a)
    {
        auto c = v[a]->cartesian();
        out[3 * a + 0] = c[0];
        out[3 * a + 1] = c[1];
        out[3 * a + 2] = c[2];
    }
}

double ref_plm(const double v[3], double theta, double sgn)
{
    Reconstruction r;
    r.setPlmTheta(theta);
    return r.reconstruct(&v[1], sgn > 0 ? Reconstruction::PLM_C2R : Reconstruction::PLM_C2L);
}

// op: 0 FD_C2R, 1 FD_C2L, 2 FV_C2R, 3 FV_C2L, 4 FV_C2A, 5 FV_A2C
// mode: 0 JiangShu96, 1 Borges08, 2 ShenZha10
double ref_weno5(const double v[5], int op, int mode, double A)
{
    static const Reconstruction::Operation ops[6] = {
        Reconstruction::WENO5_FD_C2R, Reconstruction::WENO5_FD_C2L,
        Reconstruction::WENO5_FV_C2R, Reconstruction::WENO5_FV_C2L,
        Reconstruction::WENO5_FV_C2A, Reconstruction::WENO5_FV_A2C};
    static const Reconstruction::SmoothnessIndicator modes[3] = {
        Reconstruction::OriginalJiangShu96, Reconstruction::ImprovedBorges08,
        Reconstruction::ImprovedShenZha10};
    Reconstruction r;
    r.setSmoothnessIndicator(modes[mode]);
    r.setShenZha10A(A);
    return r.reconstruct(&v[2], ops[op]);
}

int ref_quartic(const double d[5], double roots[4])
{
    QuarticPolynomial q(d[0], d[1], d[2], d[3], d[4]);
    return q.solve(roots);
}

// new_state + estimate_from_cons + Anton, then Noble, as
// RelativisticMHD::fromConserved does.
int ref_rmhd_c2p(const double U[8], double gamma, double pfloor, double P[8], int* which)
{
    ConsToPrimRMHD s;
    s.set_gamma(gamma);
    s.set_pressure_floor(pfloor);
    s.new_state(U);
    s.estimate_from_cons();
    int e;
    if ((e = s.solve_anton2dzw(P)) == 0) { *which = 0; return 0; }
    if ((e = s.solve_noble1dw(P)) == 0) { *which = 1; return 0; }
    *which = -1;
    return e;
}

}
