/*
 * mara_oracle.h -- CPU restatement of Mara2's Godunov time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * with; nothing in the product (mara2_amd/, include/) links, loads or calls
 * it.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * use it.  Every function follows the reference file:line cited beside it,
 * in the same floating-point operation order, compiled without contraction.
 *
 * Host arrays use the reference MeshData::P layout: row-major
 * (n1+2g1, n2+2g2, n3+2g3, nq), field index innermost.
 */
#ifndef MARA_ORACLE_H
#define MARA_ORACLE_H

#include "../include/mara_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* UnitVector xhat, yhat, zhat as cartesian triples (src/UnitVector.cpp:9-39). */
void oracle_unit_vectors(double out[9]);

/* Reconstruction::plm on v[0..2] centred at v[1] (src/Reconstruction.cpp:55-66). */
double oracle_plm(const double v[3], double theta, double sgn);

/* Reconstruction::weno5 on v[0..4] centred at v[2] (src/Reconstruction.cpp:68-169).
   op: 0 FD_C2R, 1 FD_C2L, 2 FV_C2R, 3 FV_C2L, 4 FV_C2A, 5 FV_A2C.
   mode: 0 JiangShu96, 1 Borges08, 2 ShenZha10. */
double oracle_weno5(const double v[5], int op, int mode, double shenzha_A);

/* QuarticPolynomial::solve (src/QuarticPolynomial.cpp:19-144). */
int oracle_quartic(const double d[5], double roots[4]);

/* ConsToPrimRMHD: new_state + estimate_from_cons + solve_anton2dzw, falling
   back to solve_noble1dw (src/ConservationLaws.cpp:627-657,
   src/ConsToPrimRMHD.cpp:41-349).  Returns 0 on success, else the Noble
   error code; *which = 0 if Anton succeeded, 1 if Noble did, -1 if none. */
int oracle_rmhd_c2p(const double U[8], double gamma, double pfloor, double P[8], int* which);

/* Both inversion solvers' status codes for U: codes[0] Anton, codes[1]
   Noble (0 success, 3 tau < 0, 5 p < 0, 7 v.v >= 1, 8 iteration limit, ...). */
void oracle_rmhd_c2p_codes(const double U[8], double gamma, double pfloor, int codes[2]);

/* Diagnostics of the last oracle_advance that failed an inversion: returns
   the 0-based RK stage (-1 when the last advance had no failure) and copies
   the first failing zone's conserved state into U (nq doubles) if U != NULL. */
int oracle_last_failure(double* U);

/* ConservationLaw::fromPrimitive for a law along unit normal n: writes
   U[nq], F[nq], A[10].  Returns MARA_OK, MARA_ERR_SUPERLUMINAL or
   MARA_ERR_BAD_EIGENVALUES. */
int oracle_from_primitive(int law, double gamma, const double* P, const double n[3],
                          double* U, double* F, double* A);

/* ConservationLaw::fromConserved -> P (only the primitive part is kept).
   Returns MARA_OK or MARA_ERR_INVERSION_FAILURE; *health = 1 if the
   pressure floor was applied. */
int oracle_from_conserved(int law, double gamma, double pfloor, const double* U,
                          double* P, int* health);

/* Intercell Godunov flux from two primitive face states: fromPrimitive on
   both, then the Riemann solver (src/RiemannSolvers.cpp:38-118). */
int oracle_riemann(int law, int solver, double gamma, const double* PL,
                   const double* PR, const double n[3], double* F);

/* CellCenteredFieldCT::correctGodunovFluxes on a flux array of shape
   (s0, s1, s2, nq, 3) in place (src/CellCenteredFieldCT.cpp:14-134). */
void oracle_ct_correct(double* F, const int s[3], int nq, int mag_index);

/* MeshOperator::divergence (src/MeshOperator.cpp:323-381) on a Cartesian
   mesh: flux (s0,s1,s2,nq,3) -> D (s0-1,s1-1,s2-1,nq). */
/* face areas (low side, axes 0-2), volume, cell lengths and centre of
   interior-relative cell (i, j, k) of the descriptor's mesh */
void oracle_cell_geometry(const mara_desc* d, int i, int j, int k, double out[10]);
/* MeshOperator::divergence over the padded array with the mesh's geometry */
void oracle_divergence_desc(const mara_desc* d, const double* F, int nq, double factor, double* D);
void oracle_divergence(const double* F, const int s[3], int nq, double factor,
                       const double dA[3], double dV, double* D);

/* CellCenteredFieldCT::monopole(B, cell) (src/CellCenteredFieldCT.cpp:159-229):
   B (s0,s1,s2,3) -> Mc (s0,s1,s2). */
void oracle_monopole_cell(const double* B, const int s[3], double* Mc);

/* Vector potential on faces -> cell B with guards, the reference's
   initialisation (src/Problems.cpp:660-667): A has shape (N1+1,N2+1,N3+1,3,3)
   sampled as MeshOperator::generate(face, scalars); writes B (N1,N2,N3,3). */
void oracle_vector_potential_to_B(const mara_desc* d, const double* A, double* B);

/* BoundaryCondition::applySimple on P (src/BoundaryConditions.cpp:8-176). */
int oracle_apply_bc(const mara_desc* d, double* P);

/* MethodOfLinesTVD::advance (src/SolutionSchemes.cpp:108-275), in place. */
int oracle_advance(const mara_desc* d, double* P, double t0, double dt, mara_report* rep);

/* FieldOperator::getCourantTimestep over the interior with
   L = linearCellDimension (src/FieldOperator.cpp:125-169). */
int oracle_courant_dt(const mara_desc* d, const double* P, double* dt);

/* max |monopole| over interior cells of the cell B held in P. */
int oracle_monopole_max(const mara_desc* d, const double* P, double* out);

/* OpenMP thread count used by oracle_advance / oracle_courant_dt
   (1 unless built with -fopenmp).  Results do not depend on it. */
void oracle_set_threads(int n);
int  oracle_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif
