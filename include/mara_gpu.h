/*
 * mara_gpu.h -- C ABI of the MI355X (gfx950) Godunov finite-volume step.
 *
 * The library advances one mesh block of Mara2's method-of-lines TVD scheme
 * on one GPU.  Every per-zone operation (reconstruction, Riemann fluxes,
 * flux-CT, divergence, Runge-Kutta stage update, primitive recovery, ghost
 * fill, CFL reduction) runs in hand-written HIP kernels; the host side only
 * validates the descriptor, owns device memory and enqueues launches.
 *
 * Conventions shared by every entry point:
 *   - Return value is a status code: MARA_OK (0) or one of MARA_ERR_*.
 *     mara_error_string() turns it into text.
 *   - Host arrays of primitive data use the reference MeshData::P layout
 *     (src/MeshData.cpp:8-28): a row-major array of shape
 *     (n1+2g1, n2+2g2, n3+2g3, nq), guard zones included, field index
 *     innermost.  Flat axes have n=1, g=0.
 *   - Field order is the reference's: density, velocity1-3, pressure,
 *     magnetic1-3 (src/ConservationLaws.cpp:6-24).
 *   - No pointer passed in is retained after the call returns.
 *   - A block handle may be used from one host thread at a time; it owns
 *     one HIP stream.  Calls are synchronous on return unless noted.
 */
#ifndef MARA_GPU_H
#define MARA_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define MARA_OK                    0
#define MARA_ERR_INVALID_ARGUMENT  1  /* bad pointer, size or enum value */
#define MARA_ERR_GUARD_TOO_SMALL   2  /* src/SolutionSchemes.cpp:45-48 logic_error */
#define MARA_ERR_UNSUPPORTED       3  /* configuration the reference cannot run */
#define MARA_ERR_HIP               4  /* a HIP runtime call failed */
#define MARA_ERR_INVERSION_FAILURE 5  /* src/FieldOperator.cpp:57-80 InversionFailure */
#define MARA_ERR_SUPERLUMINAL      6  /* src/ConservationLaws.cpp:684-687 runtime_error */
#define MARA_ERR_BAD_EIGENVALUES   7  /* src/ConservationLaws.cpp:796-833 runtime_error */
#define MARA_ERR_NO_DEVICE         8  /* no usable gfx950 device */
#define MARA_ERR_CAPACITY          9  /* request exceeds a stated capacity */

/* ---- enumerations (each names the reference class it stands for) -------- */
#define MARA_LAW_HYDRO   0  /* NewtonianHydro   src/ConservationLaws.cpp:233-343 */
#define MARA_LAW_MHD     1  /* NewtonianMHD     src/ConservationLaws.cpp:347-605 */
#define MARA_LAW_RMHD    2  /* RelativisticMHD  src/ConservationLaws.cpp:609-945 */

#define MARA_RECON_PCM   0  /* MethodOfLines     src/IntercellFluxSchemes.cpp:35-63 */
#define MARA_RECON_PLM   1  /* MethodOfLinesPlm  src/IntercellFluxSchemes.cpp:69-117 */
#define MARA_RECON_WENO5 2  /* MethodOfLinesPlm structure with Reconstruction::weno5
                               (WENO5_FD_C2R/C2L, ImprovedShenZha10, A=50),
                               src/Reconstruction.cpp:68-169 */

#define MARA_RIEMANN_HLLE 0 /* HlleRiemannSolver                src/RiemannSolvers.cpp:38-57 */
#define MARA_RIEMANN_HLLC 1 /* HllcNewtonianHydroRiemannSolver  src/RiemannSolvers.cpp:63-118 */

#define MARA_BC_PERIODIC   0 /* PeriodicBoundaryCondition   src/BoundaryConditions.cpp:24-60 */
#define MARA_BC_OUTFLOW    1 /* OutflowBoundaryCondition    src/BoundaryConditions.cpp:66-105 */
#define MARA_BC_REFLECTING 2 /* ReflectingBoundaryCondition src/BoundaryConditions.cpp:111-176 */

#define MARA_GEOM_CARTESIAN 0 /* CartesianMeshGeometry src/CartesianMeshGeometry.cpp */
#define MARA_GEOM_SPHERICAL 1 /* SphericalMeshGeometry src/SphericalMeshGeometry.cpp:
                                 axes (r, theta, phi), lower/upper their
                                 ranges, optional logarithmic radial edges */

#define MARA_PREC_FP64   0  /* every operation in IEEE binary64, no contraction */
#define MARA_PREC_MIXED  1  /* reconstruction, Riemann fluxes and flux divergence
                               in binary32; RK accumulation, flux-CT, primitive
                               recovery and the RMHD quartic in binary64 */
#define MARA_PREC_FP64_FAST 2 /* every operation in binary64, but not the
                               reference's rounding: fused multiply-add,
                               reciprocals and square roots refined from the
                               hardware approximations, axis-aligned face
                               normals; held to relative L1 <= 1e-12 of the
                               reference's conserved fields after 10 steps.
                               Hydro blocks with PLM on 2-D and 3-D meshes run
                               their own stage kernel, 2-D MHD blocks the MHD
                               tiles in this arithmetic; every other block runs
                               the MARA_PREC_FP64 kernels.  Blocks and block
                               groups; not in batches. */

/* ---- block descriptor ---------------------------------------------------- */
typedef struct mara_desc
{
    int    law;            /* MARA_LAW_* */
    double gamma;          /* ConservationLaw::setGammaLawIndex; reference default 5/3 */
    double pressure_floor; /* ConservationLaw::setPressureFloor; <= 0 disables */
    int    recon;          /* MARA_RECON_*; stencil size PCM 1, PLM 2, WENO5 3 */
    double plm_theta;      /* MethodOfLinesPlm::setPlmTheta; reference default 1.5 */
    int    riemann;        /* MARA_RIEMANN_*; HLLC requires MARA_LAW_HYDRO */
    int    bc;             /* MARA_BC_*, applied on every axis with guards */
    int    cells[3];       /* interior cells per axis (CartesianMeshGeometry shape) */
    int    guards[3];      /* MeshData boundary shape; 0 on flat axes */
    double lower[3];       /* CartesianMeshGeometry::setLowerUpper */
    double upper[3];
    int    rk_order;       /* MethodOfLinesTVD::setRungeKuttaOrder, 1..3 */
    int    field_ct;       /* 1 unless MethodOfLinesTVD::setDisableFieldCT(true) */
    int    precision;      /* MARA_PREC_* */
    double cooling_rate;   /* NewtonianMHD::setCoolingRate; > 0 (MARA_LAW_MHD
                              only) adds the auxiliary
                              ninth field and the cooling source
                              (src/ConservationLaws.cpp:354-358,489-501);
                              <= 0 (the reference default) is off */
    int    geometry;       /* MARA_GEOM_*; 0 (Cartesian) when zeroed.  The
                              spherical mesh runs the flux-array kernels with
                              the reference's face areas, volumes and cell
                              lengths (SphericalMeshGeometry::faceArea,
                              cellVolume, cellLength) and hands source terms
                              its cell centres (r, theta, phi); not with
                              flux-CT, viscous fluxes, batches, block groups
                              or mara_advance_noalloc */
    int    log_radial;     /* SphericalMeshGeometry::setUseLogarithmicRadialBinning */
} mara_desc;

/* Per-call diagnostics.  The reference throws at the first failure; the GPU
   finishes the launch, counts every failing zone and reports the first one
   (smallest linear interior index), then returns the matching status. */
typedef struct mara_report
{
    int       status;           /* same value the call returns */
    long long c2p_failures;     /* zones whose fromConserved failed (StateFailure) */
    int       first_failure[3]; /* interior-relative (i,j,k) of the first of them */
    long long floor_applied;    /* zones where the pressure floor was used */
    long long eigen_failures;   /* RMHD quartic gave nr not in {2,4} */
    long long superluminal;     /* RMHD states with v.v >= 1 */
} mara_report;

typedef struct mara_block mara_block;

/* ---- library -------------------------------------------------------------- */
const char* mara_version(void);
const char* mara_error_string(int status);

/* Number of fields of a law (getNumConserved): 5 hydro, 8 MHD/RMHD. */
int mara_num_fields(int law);

/* Number of fields of a descriptor's state (NewtonianMHD::getNumConserved,
   src/ConservationLaws.cpp:503-506): mara_num_fields(law), or 9 for the MHD
   law with a positive cooling_rate (the auxiliary field last); -1 for a
   null descriptor. */
int mara_desc_num_fields(const mara_desc* desc);

/* Stencil size of a reconstruction (IntercellFluxScheme::getStencilSize). */
int mara_stencil_size(int recon);

/* ---- one mesh block --------------------------------------------------------- */

/* Validate the descriptor and allocate the device state of one block.
   Replaces constructing MeshData (src/MeshData.cpp:8) plus a
   MethodOfLinesTVD wired with its collaborators (src/Problems.cpp:159-174). */
int mara_block_create(const mara_desc* desc, mara_block** out);
int mara_block_destroy(mara_block* block);

/* Copy primitives, guards included, in MeshData::P layout (see above),
   between host and device.  Count of doubles = prod(n+2g) * nq.  The
   download's guards hold the boundary condition of its interior, as after
   the reference's advance (a periodic block on the fast tiles fills them
   here rather than after every stage). */
int mara_block_upload_prim(mara_block* block, const double* P);
int mara_block_download_prim(mara_block* block, double* P);

/* One full Runge-Kutta step of size dt.  Replaces
   MethodOfLinesTVD::advance(MeshData&, double t0, double dt)
   (src/SolutionSchemes.cpp:108-275).  report may be NULL. */
int mara_block_advance(mara_block* block, double t0, double dt, mara_report* report);

/* min over interior zones of L / max(lambda_x, lambda_y, lambda_z), no CFL
   factor.  Replaces FieldOperator::getCourantTimestep(P, L) with
   L = MeshOperator::linearCellDimension() (src/FieldOperator.cpp:125-169,
   src/MeshOperator.cpp:292-307). */
int mara_block_courant_dt(mara_block* block, double* dt);

/* Fill guard zones of the resident primitives.  Replaces
   MeshData::applyBoundaryCondition -> BoundaryCondition::applySimple
   (src/MeshData.cpp:148-151, src/BoundaryConditions.cpp:8-18). */
int mara_block_apply_bc(mara_block* block);

/* max |M| over interior zones of CellCenteredFieldCT::monopole(B, cell)
   (src/CellCenteredFieldCT.cpp:159-229), B = cell-centred field with
   guards.  MARA_ERR_UNSUPPORTED for laws without a magnetic field. */
int mara_block_monopole_max(mara_block* block, double* max_abs);

/* The maraMainLoop contract (src/Mara.cpp:42-88) without a host round trip
   per step: up to max_steps steps of dt = min(cfl * courant_dt, t_final - t)
   while t < t_final.  The time step, the clock and the stop condition live
   on the device; the host waits once at the end.  On return *t holds the
   new time and *steps the number of steps taken. */
int mara_block_run(mara_block* block, double cfl, double t_final, int max_steps,
                   double* t, int* steps, mara_report* report);

/* Source terms: the counterpart of MethodOfLinesTVD::setSourceTermsFunction
   (src/SolutionSchemes.hpp:36, applied in src/SolutionSchemes.cpp:203-217).
   The per-zone function is device code in a gfx950 code object written with
   include/mara_source.h; the block loads `code_object`, takes its kernel
   `kernel`, and before each RK stage evaluates S at every interior zone from
   the stage's primitives, then adds S to the flux divergence L in the zone
   update.  params (up to 16) reach the function as given; code_object = NULL
   removes the source.  A block of the fused hydro path switches to the
   flux-array path while a source is set (its fused update has no source
   term).  MARA_ERR_INVALID_ARGUMENT when the file or the kernel cannot be
   loaded.  Batches take no sources. */
#define MARA_SOURCE_MAX_PARAMS 16
int mara_block_set_source(mara_block* block, const char* code_object, const char* kernel, const double* params,
                          int n_params);
/* The 8 auxiliary values passed to the source function (the reference
   passes its star-particle data there); they start at zero. */
int mara_block_set_source_aux(mara_block* block, const double* t);

/* Viscous fluxes (MethodOfLinesTVD::setViscousFluxFunction,
   src/SolutionSchemes.cpp:91-94 and :183-186): a per-face function in a
   gfx950 code object (written with include/mara_source.h,
   MARA_VISCOUS_FLUX_KERNEL), run on every stage after the Godunov fluxes and
   before the divergence, with up to MARA_SOURCE_MAX_PARAMS parameters and the
   auxiliary data of mara_block_set_source_aux.  The block takes the
   flux-array path while it is set.  A null code_object removes it.  With
   flux-CT the function sees the CT-corrected fluxes, as the reference's
   fluxCorrection hands them over; in the mixed form it sees binary64 copies
   of the binary32 fluxes, and the update reads their binary32 rounding.
   MARA_ERR_UNSUPPORTED on a spherical mesh; MARA_ERR_INVALID_ARGUMENT for a
   file or kernel that cannot be loaded. */
int mara_block_set_viscous_flux(mara_block* block, const char* code_object, const char* kernel,
                                const double* params, int n_params);

/* mara_block_run replays its steps from a HIP graph of 16 steps, captured on
   the first run (enable=1, the default), or launches every kernel from the
   host (enable=0).  Results are the same either way; runs with stage timing
   enabled launch from the host.  No counterpart in the reference: the
   reference steps on the host. */
int mara_block_set_graphs(mara_block* block, int enable);

/* Kernel timing on the block's stream: a pair of HIP events around each
   launch of the stage kernels (the fused kernel, or the flux and update
   kernels), recorded without host waits, for up to 512 stages.
   enable=1 starts a fresh tally. */
int mara_block_set_timing(mara_block* block, int enable);
/* Average duration in milliseconds and launch count of the stage kernels
   since timing was enabled (waits for the last recorded event); *name
   receives a static string naming the kernel. */
int mara_block_stage_kernel_time(mara_block* block, double* avg_ms, long long* launches,
                                 const char** name);

/* ---------------------------------------------------------------------------
   Launch shapes and the autotuner (no counterpart in the reference, which
   has no kernels).  A launch shape divides a stage's zones among
   work-groups differently; it never changes the arithmetic, so every shape
   gives bit-identical results.
   --------------------------------------------------------------------------- */
typedef struct mara_shape
{
    int pencil;  /* 1-D stages: faces per pencil work-group, 128, 256 or 512
                    (the relativistic law runs 256 whatever is set) */
    int zplanes; /* fused 3-D hydro stages: z planes a work-group marches,
                    1 to 1024 (the autotuner tries 4, 8, 16, 32 and the
                    counts that make the tile count a whole number of
                    device-filling rounds) */
    int rows;    /* flux-array stages: work-group rows of 64 zones; fused
                    2-D MHD stages: tiles of 64 x (4 * rows) zones; 1, 2 or 4 */
} mara_shape;

/* The block's launch shape.  A new block starts from the tuning table's
   entry for its descriptor, else from pencil 256, zplanes 16 and rows 4
   (rows 2 for a fused 2-D MHD block).
   set: MARA_ERR_INVALID_ARGUMENT for a value outside the lists above. */
int mara_block_get_shape(const mara_block* block, mara_shape* shape);
int mara_block_set_shape(mara_block* block, const mara_shape* shape);

/* Autotuner: times the candidates of the shape the block's stages use
   (pencil for 1-D blocks of two or three stages, zplanes for fused 3-D
   hydro blocks, rows for the flux-array path and the fused 2-D MHD tiles;
   nothing to choose for the fused 2-D hydro blocks): `repeats` launches of
   a first RK stage per candidate after one untimed launch, candidates in
   turn over three rounds, each candidate's fastest round counting; on the
   block's own state with dt = 0 and a scratch control block, so the state,
   clock and counters are untouched.  Sets the fastest candidate on the
   block and records it in the process's tuning table under the block's
   descriptor.  *chosen (optional) receives the shape, *best_ms (optional)
   the fastest mean stage time in ms. */
int mara_block_autotune(mara_block* block, int repeats, mara_shape* chosen, double* best_ms);

/* The process's tuning table: look up the shape recorded for a descriptor
   (returns 1 and fills *shape when found, else 0), write every entry to a
   text file, or merge the entries of such a file. */
int mara_tune_lookup(const mara_desc* desc, mara_shape* shape);
int mara_tune_save(const char* path);
int mara_tune_load(const char* path);

/* Bytes of device memory the block holds (state, fluxes, scratch). */
long long mara_block_device_bytes(const mara_block* block);

/* Device bytes a block with this descriptor would hold, without allocating
   (used for admission by HBM budget). */
long long mara_desc_device_bytes(const mara_desc* desc);

/* ---------------------------------------------------------------------------
   One step with no device allocation.  No counterpart in the reference: it
   serves callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc).  The same
   arithmetic as mara_block_advance on a block created from desc and loaded
   with P_in, with every working buffer (the state in the device layout, its
   stage copies, the RK state, flux arrays where the path needs them, the
   control block and the CFL bins) carved from one statically sized device
   array of MARA_SCRATCH_BYTES per GPU, which the library holds from its
   first use on that GPU until the process ends.  P_in and P_out are device
   pointers to primitives in MeshData::P layout (guards included); they may
   be the same array.  stream is a hipStream_t (NULL: the null stream); the
   call waits for it before returning.  MARA_ERR_CAPACITY when the
   descriptor's buffers exceed MARA_SCRATCH_BYTES.  Every call on one GPU
   shares that GPU's array: calls that overlap on one GPU (other host
   threads, other streams) race and must not be made.  The launch shape is
   the default one (the tuning table is not consulted).
   --------------------------------------------------------------------------- */
#define MARA_SCRATCH_BYTES (32ll * 1024 * 1024)
int mara_advance_noalloc(const mara_desc* desc, const double* P_in, double* P_out, double t0, double dt,
                         void* stream, mara_report* report);

/* ---------------------------------------------------------------------------
   Batches: many independent solves resident in one pool of device memory,
   advanced together with one launch per kernel type and solve class per RK
   stage.  Each solve keeps the single block's arithmetic, its own time step,
   clock, stop condition and failure counters, so its result is bitwise the
   same alone or in any mix.  A solve is the maraMainLoop contract of
   mara_block_run (src/Mara.cpp:42-88) on one MeshData: dt = min(cfl *
   courant_dt, t_final - t) while t < t_final, at most max_steps steps.
   The reference runs such solves one process at a time; this has no single
   counterpart there and serves callers that run many small problems (the
   serving loop, parameter sweeps).
   --------------------------------------------------------------------------- */
typedef struct mara_batch mara_batch;

/* A batch owning hbm_budget bytes of device memory for solve state and up
   to max_solves solves; rk_order and precision are shared by its solves.
   MARA_ERR_CAPACITY if the budget cannot be allocated. */
int mara_batch_create(long long hbm_budget, int rk_order, int precision, int max_solves,
                      mara_batch** out);
int mara_batch_destroy(mara_batch* batch);

/* Device bytes one solve with this descriptor takes from the pool. */
long long mara_batch_item_bytes(const mara_desc* desc, int precision);
/* Pool bytes not held by a solve (the largest admissible solve may be
   smaller: the pool is first-fit). */
long long mara_batch_free_bytes(const mara_batch* batch);

/* Admit a solve: P in MeshData::P layout with guards filled, as
   mara_block_upload_prim; the clock starts at 0.  *slot receives its handle.
   MARA_ERR_CAPACITY when no slot or no contiguous pool range is free (the
   caller retries after a retire); MARA_ERR_UNSUPPORTED when desc's rk_order
   or precision differ from the batch's. */
int mara_batch_admit(mara_batch* batch, const mara_desc* desc, const double* P, double cfl,
                     double t_final, int max_steps, int* slot);

/* Advance every unfinished solve by up to n_steps of its own steps; returns
   after the device has drained.  1-D solves of an RK2 batch run on a second
   stream and may take up to 32 * n_steps steps in one call (more while large
   solves make the others' steps slow), so that they are not held to the pace
   of the large ones; a solve's result does not depend on how its steps are
   grouped into calls.  *n_active receives the number of admitted solves not
   yet finished. */
int mara_batch_step(mara_batch* batch, int n_steps, int* n_active);

/* With enable=1 (the default) mara_batch_step launches one round of kernels
   (a step of the large solves and the small 1-D solves' steps beside it) as a
   HIP graph, captured when the set of solves changed and replayed n_steps
   times; with enable=0 it launches every kernel from the host.  Results are
   the same either way.  No counterpart in the reference. */
int mara_batch_set_graphs(mara_batch* batch, int enable);

/* Source terms for every solve of a batch (the batch form of
   mara_block_set_source, MethodOfLinesTVD::setSourceTermsFunction,
   src/SolutionSchemes.cpp:203-217): the code object's kernel <kernel>_batch,
   which MARA_SOURCE_KERNEL defines beside <kernel> (include/mara_source.h),
   evaluates the function at every interior zone of every solve before each
   stage, one launch for all of them; the update adds S to L.  Set before
   the first admission (MARA_ERR_INVALID_ARGUMENT once solves are resident);
   solves admitted with a source take the flux-array kernels.  A null
   code_object removes it.  MARA_ERR_UNSUPPORTED at admission for a cooled
   NewtonianMHD solve, as for a single block. */
int mara_batch_set_source(mara_batch* batch, const char* code_object, const char* kernel, const double* params,
                          int n_params);
/* The 8 auxiliary values passed to the batch's source function. */
int mara_batch_set_source_aux(mara_batch* batch, const double* t);

/* With workgroups_per_cu > 0 (at most 16) the 2-D and 3-D hydro + PLM
   tiles of a stage run as a work queue: a fixed grid of workgroups_per_cu
   work-groups per CU that claim the same tiles in the same per-XCD order
   from atomic counters, instead of one work-group per tile.  Results are
   unchanged; 0 (the default) launches one work-group per tile.  No
   counterpart in the reference (a persistent-kernel experiment for the
   serving batch). */
int mara_batch_set_persistent(mara_batch* batch, int workgroups_per_cu);

/* The most solves the express lane may hold (default 0: every 2-D and 3-D
   solve stays on the tile lane).  Solves already there stay. */
int mara_batch_set_express(mara_batch* batch, int max_solves);

/* *lane receives the lane a solve runs on: 0 the tile lane (2-D and 3-D solves), 1 the
   pencil lane (1-D solves of an RK2 batch, up to 32 of their steps per
   round on a second stream), 2 the express lane (a 2-D or 3-D solve whose
   remaining steps exceed 4x the median of the running tile-lane solves and
   1000, moved at the start of a mara_batch_step and kept there, up to 32 of
   its steps per round on a third stream; see mara_batch_set_express).  A solve's
   result does not depend on its lane.  MARA_ERR_INVALID_ARGUMENT for a
   free slot. */
int mara_batch_lane(const mara_batch* batch, int slot, int* lane);

/* Launch shape of the batch's pencil tiles (1-D solves), fused 3-D tiles
   (zplanes) and fused 2-D MHD tiles (rows); at create pencil 128, zplanes
   16, rows 2.  Takes effect at the next step. */
int mara_batch_get_shape(const mara_batch* batch, mara_shape* shape);
int mara_batch_set_shape(mara_batch* batch, const mara_shape* shape);

/* State of one solve as of the last mara_batch_step: *done is 1 once it
   reached t_final, its step limit or a failure.  Returns the solve's status
   (the codes of mara_block_run) and fills report when non-null. */
int mara_batch_query(mara_batch* batch, int slot, int* done, double* t, int* steps,
                     mara_report* report);

/* Slots whose solve finished as of the last mara_batch_step and is not yet
   retired, in slot order: up to capacity of them into slots, their number
   into *count. */
int mara_batch_finished(const mara_batch* batch, int* slots, int capacity, int* count);

/* Copy the solve's primitives into P (MeshData::P layout; skipped when P
   is null) and release its slot and memory. */
int mara_batch_retire(mara_batch* batch, int slot, double* P);

/* ---------------------------------------------------------------------------
   Block groups: one mesh split into blocks on one GPU, the single-device half
   of BlockDecomposition (src/BlockDecomposition.cpp).  Each block runs its
   stages on a stream of its own, so the blocks' kernels overlap; once all of
   a stage is done, the group's stream fills their guard zones from the
   neighbours' edge layers axis by axis, left then right, and
   applies the physical condition at the mesh's walls, as BlockDecomposedBC
   under BoundaryCondition::applySimple does (src/BlockDecomposition.cpp:11-92,
   src/BoundaryConditions.cpp:8-18).  Every block keeps the whole mesh's cell
   sizes, so a group's step is the single block's step zone for zone.
   --------------------------------------------------------------------------- */
typedef struct mara_group mara_group;

#define MARA_GROUP_MAX_BLOCKS 1024

/* Partition `which` of n elements over p partitions: its element count and
   first element.  BlockDecomposition::partition and ::startIndex
   (src/BlockDecomposition.cpp:187-202): the first n % p partitions hold one
   element more. */
int mara_partition(int n, int p, int which, int* count, int* start);

/* Split desc's mesh into dims[0] x dims[1] x dims[2] blocks (block
   coordinates row-major, axis 2 fastest, as a Cartesian communicator ranks
   them; BlockDecomposition::getPatchRegion, src/BlockDecomposition.cpp:162-175).
   MARA_ERR_UNSUPPORTED for MHD and RMHD with flux-CT (the face fluxes of a
   neighbour's edge are needed inside a stage) and for a split that leaves a
   block fewer than two cells or fewer cells than guards along an axis;
   MARA_ERR_CAPACITY beyond MARA_GROUP_MAX_BLOCKS blocks. */
int mara_group_create(const mara_desc* desc, const int* dims, mara_group** out);
int mara_group_destroy(mara_group* group);
int mara_group_num_blocks(const mara_group* group);
/* First interior cell and cell counts of block `index` in the mesh. */
int mara_group_block_region(const mara_group* group, int index, int* start, int* cells);
long long mara_group_device_bytes(const mara_group* group);

/* Primitives of the whole mesh in MeshData::P layout, guards included.  The
   upload gives each block its window of P (guards included); the download
   writes every block's interior and the mesh's own guard zones. */
int mara_group_upload_prim(mara_group* group, const double* P);
int mara_group_download_prim(mara_group* group, double* P);

/* The guard exchange on the resident primitives (MeshData::applyBoundaryCondition
   with a BlockDecomposedBC, src/MeshData.cpp:148-151). */
int mara_group_apply_bc(mara_group* group);

/* One Runge-Kutta step of size dt of every block, the guard exchange after
   each stage (MethodOfLinesTVD::advance, src/SolutionSchemes.cpp:108-275,
   with a BlockDecomposedBC).  The report sums the blocks' counts; its first
   failure is the one with the smallest index of the whole mesh, in the
   mesh's interior coordinates. */
int mara_group_advance(mara_group* group, double t0, double dt, mara_report* report);

/* Minimum over the blocks of mara_block_courant_dt: the whole mesh's CFL
   time step. */
int mara_group_courant_dt(mara_group* group, double* dt);

#ifdef __cplusplus
}
#endif

#endif /* MARA_GPU_H */
