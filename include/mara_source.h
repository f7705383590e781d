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
} mara_source_args;

/* MARA_SOURCE_KERNEL(name, fn): an extern "C" kernel `name` that evaluates
   fn(x, y, z, t, P, S, params) at every interior zone, one thread per zone,
   256 threads per work-group. */
#define MARA_SOURCE_KERNEL(name, fn)                                                              \
    extern "C" __global__ void __launch_bounds__(256) name(mara_source_args a)                    \
    {                                                                                             \
        const long long z = (long long)blockIdx.x * blockDim.x + threadIdx.x;                      \
        const long long nx = a.n[0], nxy = (long long)a.n[0] * a.n[1];                             \
        if (z >= nxy * a.n[2]) return;                                                            \
        const int ic = (int)(z % nx), jc = (int)((z / nx) % a.n[1]), kc = (int)(z / nxy);          \
        const long long c = ((long long)(kc + a.g[2]) * a.N[1] + (jc + a.g[1])) * a.N[0] + (ic + a.g[0]); \
        double P[8], S[8];                                                                        \
        for (int q = 0; q < 8; ++q)                                                               \
        {                                                                                         \
            P[q] = q < a.nq ? a.P[q * a.fs + c] : 0.0;                                            \
            S[q] = 0.0;                                                                           \
        }                                                                                         \
        const double x = a.lower[0] + a.dx[0] * (ic + 0.5);                                       \
        const double y = a.lower[1] + a.dx[1] * (jc + 0.5);                                       \
        const double zz = a.lower[2] + a.dx[2] * (kc + 0.5);                                      \
        fn(x, y, zz, a.t, P, S, a.params);                                                        \
        for (int q = 0; q < a.nq; ++q) a.S[q * a.fs + c] = S[q];                                  \
    }

#endif /* MARA_SOURCE_H */
