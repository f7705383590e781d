/* mara_source.h -- writing a source-term function for mara_block_set_source.

   The reference takes a sourceTermsFunction, a host callable
   S = f(x, y, z, t[8], P) per zone (src/Mara.hpp:67, used by
   MethodOfLinesTVD::advance, src/SolutionSchemes.cpp:203-217).  On the GPU the
   same per-zone function is device code: write it as a __device__ function
   with the signature below, wrap it with MARA_SOURCE_KERNEL, and compile the
   file to a gfx950 code object

       hipcc --genco --offload-arch=gfx950 -O3 -ffp-contract=off \
             -I<repo>/include my_source.hip -o my_source.hsaco

   (-ffp-contract=off keeps the arithmetic the reference's).  The block loads
   the code object with mara_block_set_source and, before each RK stage,
   evaluates S at every interior zone from the stage's primitives; the zone
   update adds S to the flux divergence L, as the reference does.

   The function receives the zone's cell-centre coordinates (the reference's
   coordinateAtIndex), the 8 auxiliary values set by
   mara_block_set_source_aux (the reference passes its star-particle data
   there), the zone's primitives P[0..nq) and the parameters given to
   mara_block_set_source; it writes S[0..nq) (S starts zeroed). */
#ifndef MARA_SOURCE_H
#define MARA_SOURCE_H

#include <hip/hip_runtime.h>

#include "mara_gpu.h" /* MARA_SOURCE_MAX_PARAMS */

/* The launch record the block passes to a source kernel (by value). */
typedef struct mara_source_args
{
    const double* P;   /* stage primitives, field q of cell c at P[q * fs + c] */
    double* S;         /* source terms, same layout */
    long long fs;      /* field stride */
    int N[3];          /* cells including guards; c = (k * N[1] + j) * N[0] + i */
    int g[3];          /* guard zones */
    int n[3];          /* interior cells */
    int nq;            /* fields of the law: 5 hydro, 8 MHD / RMHD */
    int pad;
    double lower[3];   /* domain lower corner */
    double dx[3];      /* cell size: (upper - lower) / n */
    double t[8];       /* auxiliary data (mara_block_set_source_aux) */
    double params[MARA_SOURCE_MAX_PARAMS];
    const double* centre[3]; /* null on a Cartesian mesh; on a spherical one
                                the n[a] cell-centre coordinates of each axis
                                (SphericalMeshGeometry::coordinateAtIndex) */
} mara_source_args;

/* MARA_SOURCE_KERNEL(name, fn): an extern "C" kernel `name` that evaluates
   fn(x, y, z, t, P, S, params) at every interior zone, one thread per zone,
   256 threads per work-group; and a kernel `name_batch` (for
   mara_batch_set_source) that does the same for solve blockIdx.y of an
   array of launch records. */
#define MARA_SOURCE_ZONE(a, fn)                                                                   \
    do                                                                                            \
    {                                                                                             \
        const long long z = (long long)blockIdx.x * blockDim.x + threadIdx.x;                      \
        const long long nx = (a).n[0], nxy = (long long)(a).n[0] * (a).n[1];                       \
        if (z >= nxy * (a).n[2]) break;                                                           \
        const int ic = (int)(z % nx), jc = (int)((z / nx) % (a).n[1]), kc = (int)(z / nxy);        \
        const long long c = ((long long)(kc + (a).g[2]) * (a).N[1] + (jc + (a).g[1])) * (a).N[0]   \
                            + (ic + (a).g[0]);                                                    \
        double P[8], S[8];                                                                        \
        for (int q = 0; q < 8; ++q)                                                               \
        {                                                                                         \
            P[q] = q < (a).nq ? (a).P[q * (a).fs + c] : 0.0;                                      \
            S[q] = 0.0;                                                                           \
        }                                                                                         \
        const double x = (a).centre[0] ? (a).centre[0][ic] : (a).lower[0] + (a).dx[0] * (ic + 0.5); \
        const double y = (a).centre[1] ? (a).centre[1][jc] : (a).lower[1] + (a).dx[1] * (jc + 0.5); \
        const double zz = (a).centre[2] ? (a).centre[2][kc] : (a).lower[2] + (a).dx[2] * (kc + 0.5); \
        fn(x, y, zz, (a).t, P, S, (a).params);                                                    \
        for (int q = 0; q < (a).nq; ++q) (a).S[q * (a).fs + c] = S[q];                            \
    } while (0)

#define MARA_SOURCE_KERNEL(name, fn)                                                              \
    extern "C" __global__ void __launch_bounds__(256) name(mara_source_args a)                    \
    {                                                                                             \
        MARA_SOURCE_ZONE(a, fn);                                                                  \
    }                                                                                             \
    extern "C" __global__ void __launch_bounds__(256) name##_batch(const mara_source_args* items) \
    {                                                                                             \
        const mara_source_args& b = items[blockIdx.y];                                           \
        MARA_SOURCE_ZONE(b, fn);                                                                  \
    }

/* ---------------------------------------------------------------------------
   Viscous fluxes for mara_block_set_viscous_flux.  The reference takes a
   viscousFlux function (P, fluxes, T) that adds to the Godunov fluxes after
   flux-CT and before the divergence (MethodOfLinesTVD::setViscousFluxFunction,
   src/SolutionSchemes.cpp:91-94, called at :183-186).  On the GPU it is a
   per-face __device__ function
       void fn (const mara_flux_args& a, int axis, int i, int j, int k)
   that updates the fluxes of the face below cell (i, j, k) along `axis`
   (between cells (i, j, k) - e_axis and (i, j, k)), wrapped with
   MARA_VISCOUS_FLUX_KERNEL and compiled to a code object as above.
   --------------------------------------------------------------------------- */
typedef struct mara_flux_args
{
    const double* P;   /* stage primitives, field q of cell c at P[q * fs + c] */
    double* F[3];      /* Godunov fluxes along each axis (null for a flat
                          axis): field q of face f at F[a][q * ffs + f],
                          f = (k * (N[1] + 1) + j) * (N[0] + 1) + i */
    long long fs;      /* cell field stride */
    long long ffs;     /* face field stride */
    int N[3];          /* cells including guards */
    int g[3];          /* guard zones */
    int n[3];          /* interior cells */
    int nq;            /* fields of the law */
    int pad;
    double lower[3];   /* domain lower corner */
    double dx[3];      /* cell size */
    double t[8];       /* auxiliary data (mara_block_set_source_aux) */
    double params[MARA_SOURCE_MAX_PARAMS];
} mara_flux_args;

/* MARA_VISCOUS_FLUX_KERNEL(name, fn): an extern "C" kernel `name` that calls
   fn(a, axis, i, j, k) once for every face the zone update reads: along each
   axis with fluxes, faces g .. g + n of that axis on the interior rows and
   columns of the others; one thread per face, 256 per work-group, the axes
   in order x, y, z. */
#define MARA_VISCOUS_FLUX_KERNEL(name, fn)                                                        \
    extern "C" __global__ void __launch_bounds__(256) name(mara_flux_args a)                      \
    {                                                                                             \
        long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;                            \
        for (int ax = 0; ax < 3; ++ax)                                                            \
        {                                                                                         \
            if (!a.F[ax]) continue;                                                               \
            const int m0 = a.n[0] + (ax == 0), m1 = a.n[1] + (ax == 1), m2 = a.n[2] + (ax == 2);   \
            const long long cnt = (long long)m0 * m1 * m2;                                        \
            if (u < cnt)                                                                          \
            {                                                                                     \
                const int i = a.g[0] + (int)(u % m0);                                             \
                const int j = a.g[1] + (int)((u / m0) % m1);                                      \
                const int k = a.g[2] + (int)(u / ((long long)m0 * m1));                           \
                fn(a, ax, i, j, k);                                                               \
                return;                                                                           \
            }                                                                                     \
            u -= cnt;                                                                             \
        }                                                                                         \
    }

#endif /* MARA_SOURCE_H */
