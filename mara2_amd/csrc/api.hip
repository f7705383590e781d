// api.hip -- host side of the C ABI declared in include/mara_gpu.h.
//
// Owns the device state of one mesh block and enqueues the per-stage kernels
// on the block's stream (stage.hpp picks them: the fused hydro kernel of
// fused.hpp, the fused 2-D MHD kernel of fused_mhd.hpp, the 1-D pencils or
// the flux-array kernels of kernels.hpp), replays device-driven steps from a
// captured HIP graph, runs the launch-shape autotuner and loads source-term
// code objects.  Device layout and the stage structure are described in
// kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/mara_gpu.h"
#include "../../include/mara_source.h"
#include "kernels.hpp"
#include "fused.hpp"
#include "host_common.hpp"
#include "stage.hpp"

using namespace mara;
using namespace mara::host;

namespace {

constexpr int kThreads = 256;

unsigned grid_for(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

#define HIP_TRY(call) do { if (!hip_ok(call)) return MARA_ERR_HIP; } while (0)

struct Sizes
{
    long long ncell, fs, nface, ffs, real;
    int nq;
    bool ct;
    int fleshed[3];
};

Sizes sizes_of(const mara_desc* d)
{
    Sizes s;
    int N[3];
    for (int a = 0; a < 3; ++a)
    {
        N[a] = d->cells[a] + 2 * d->guards[a];
        s.fleshed[a] = N[a] > 1;
    }
    s.nq = law_nq(law_code(*d));
    s.ncell = (long long)N[0] * N[1] * N[2];
    s.fs = round_up(s.ncell, 64);
    s.nface = (long long)(N[0] + 1) * (N[1] + 1) * (N[2] + 1);
    s.ffs = round_up(s.nface, 64);
    s.real = d->precision == MARA_PREC_MIXED ? 4 : 8;
    s.ct = d->law != MARA_LAW_HYDRO && d->field_ct;
    return s;
}

long long device_bytes(const mara_desc* d, const Sizes& s)
{
    // fused and pencil blocks compute each stage's fluxes inside the stage
    // pass (no flux arrays) and step Pa -> Pb (-> Pc) -> Pa
    const bool in_pass = fused_supported(d) || pencil_supported(d) || mhd2d_supported(d);
    long long b = 3ll * s.nq * s.fs * 8;  // Pa, Pb, U
    if (in_pass && d->rk_order == 3) b += 1ll * s.nq * s.fs * 8;  // Pc
    if (!in_pass)
        b += (long long)(s.fleshed[0] + s.fleshed[1] + s.fleshed[2]) * s.nq * s.ffs * s.real;
    if (s.nq == 9) b += 1ll * s.nq * s.fs * 8;  // cooling terms
    b += 256 + CBINS * CBIN_STRIDE * 8;  // control block and CFL bins
    return b;
}

} // namespace

struct mara_block
{
    mara_desc d;
    Geom G;
    Sizes S;
    int ng;
    bool fused;
    hipStream_t stream = nullptr;
    double* Pa = nullptr;
    double* Pb = nullptr;
    double* Pc = nullptr;  // third stage buffer of the fused RK3 schedule
    double* U = nullptr;
    void* F[3] = {nullptr, nullptr, nullptr};
    void* Fv[3] = {nullptr, nullptr, nullptr};  // binary64 flux staging of a viscous hook (mixed form or flux-CT)
    int pencil = 0;  // 1-D pencil tile length (pencil_supported), else 0
    bool mhd2d = false;  // fused 2-D MHD tiles (mhd2d_supported)
    int kz = FUSED_KZ;           // fused 3-D kernels: z planes per work-group
    int rows = ROWS_DEFAULT;     // flux-array kernels: work-group rows
    int pencil_shape = PENCIL_DEFAULT;  // requested pencil length (all laws)
    bool in_group = false;    // a group member: the group's exchange fills its guards
    // the guards of Pa may not hold the ghost fill of its interior: the
    // stages of the PLM tiles of a periodic block neither read nor fill them
    // (StageArgs::wrap); whatever reads them next fills them first (fill_guards)
    bool guards_stale = false;
    double* sph = nullptr;    // spherical mesh: the geometry tables (G.sph)
    Ctl* ctl = nullptr;
    Ctl* ctl_host = nullptr;
    unsigned long long* cbins = nullptr;
    long long bytes = 0;
    // stage-kernel timing: event pairs recorded around each stage launch with
    // no host wait in between (the GPU stays busy), summed when read
    static constexpr int kTimingPairs = 512;
    bool timing = false;
    hipEvent_t ev[2 * kTimingPairs] = {};
    int ev_used = 0;
    double timed_ms = 0.0;
    long long timed_launches = 0;
    // HIP graph of kGraphSteps device-driven steps (control kernels and stage
    // launches), captured by the first mara_block_run that may use it and
    // replayed by later ones; the step's buffers and arguments never change
    // over the block's life, and steps past the stop condition are no-ops
    static constexpr int kGraphSteps = 16;
    bool graphs = true;
    hipGraphExec_t steps_exec = nullptr;
    // source terms (mara_block_set_source): the loaded code object, its
    // kernel, the launch record and the S array (field layout of P)
    hipModule_t src_module = nullptr;
    hipFunction_t src_fn = nullptr;
    mara_source_args src_args = {};
    double* Ssrc = nullptr;
    // viscous fluxes (mara_block_set_viscous_flux): code object, kernel,
    // launch record and face count; while set, the stages take the
    // flux-array path
    hipModule_t vis_module = nullptr;
    hipFunction_t vis_fn = nullptr;
    mara_flux_args vis_args = {};
    long long vis_faces = 0;
};

namespace {

// the viscous-flux kernel of the block, on the stage's primitives and the
// flux arrays the sweeps just wrote (StageArgs::after_flux)
void launch_viscous(void* ctx)
{
    mara_block* b = static_cast<mara_block*>(ctx);
    void* args[] = {&b->vis_args};
    (void)hipModuleLaunchKernel(b->vis_fn, (unsigned)((b->vis_faces + 255) / 256), 1, 1, 256, 1, 1, 0, b->stream,
                                args, nullptr);
}

void launch_bc(mara_block* b, double* P)
{
    const Geom& G = b->G;
    const long long total = 2ll * G.g[0] * G.N[1] * G.N[2] + (long long)G.n[0] * 2 * G.g[1] * G.N[2]
                            + (long long)G.n[0] * G.n[1] * 2 * G.g[2];
    if (total > 0) k_ghost<<<grid_for(total), kThreads, 0, b->stream>>>(P, G, total, b->ctl);
}

// The ghost fill of Pa where its guards may be stale, whatever the state of
// the device loop (a run's stop condition leaves the control block inactive)
void fill_guards(mara_block* b)
{
    if (!b->guards_stale) return;
    const Geom& G = b->G;
    const long long total = 2ll * G.g[0] * G.N[1] * G.N[2] + (long long)G.n[0] * 2 * G.g[1] * G.N[2]
                            + (long long)G.n[0] * G.n[1] * 2 * G.g[2];
    if (total > 0) k_ghost<<<grid_for(total), kThreads, 0, b->stream>>>(b->Pa, G, total, nullptr);
    b->guards_stale = false;
}

// whether the stage kernels of the block read their halos from the interior
// (the fast tiles and the bit-exact PLM tile of a periodic single block,
// StageArgs::wrap)
bool stage_wraps(const mara_block* b)
{
    const bool mixed = b->d.precision == MARA_PREC_MIXED;
    const bool fast_tile = b->fused && fast64_supported(&b->d) && (mixed || b->d.precision == MARA_PREC_FP64_FAST);
    const bool exact_plm = b->fused && b->d.precision == MARA_PREC_FP64 && b->d.recon == MARA_RECON_PLM;
    // every guard's periodic image is an interior zone: validate() rejects
    // axes with fewer cells than guard layers
    return (fast_tile || exact_plm) && b->G.bc == 0 && !b->in_group;
}

// The arithmetic of the CFL time step the block's last stage produces
// (the values of fast64::CflArith: 0 the reference's binary64, 1 binary32
// wave speeds, 2 and 3 the fast tiles' binary64 and binary32).
int cfl_arith(const mara_block* b)
{
    const bool fast_tile = b->fused && fast64_supported(&b->d);
    if (b->d.precision == MARA_PREC_MIXED) return fast_tile ? 3 : (law_code(b->d) < 2 ? 1 : 0);
    if (b->d.precision == MARA_PREC_FP64_FAST && fast_tile) return 2;
    if (b->d.precision == MARA_PREC_FP64_FAST && b->mhd2d) return 4;
    return 0;
}

void launch_courant(mara_block* b)
{
    const Geom& G = b->G;
    const long long total = (long long)G.n[0] * G.n[1] * G.n[2];
    long long blocks = (total + kThreads - 1) / kThreads;
    if (blocks > 2048) blocks = 2048;  // 8 work-groups per CU, grid-stride beyond
    const unsigned nb = (unsigned)blocks;
    // the first step of a run takes its dt in the arithmetic of the later
    // ones, which the form's last stage computes
    if (const int arith = cfl_arith(b))
    {
        courant_kernels_form(b->Pa, G, G.nq == 9 ? 3 : b->d.law, b->ctl, b->stream, arith, nb);
        return;
    }
    switch (b->d.law)
    {
        case 0: k_courant<0><<<nb, kThreads, 0, b->stream>>>(b->Pa, G, b->ctl); break;
        case 1:
            if (G.nq == 9) k_courant<3><<<nb, kThreads, 0, b->stream>>>(b->Pa, G, b->ctl);
            else k_courant<1><<<nb, kThreads, 0, b->stream>>>(b->Pa, G, b->ctl);
            break;
        default: k_courant<2><<<nb, kThreads, 0, b->stream>>>(b->Pa, G, b->ctl); break;
    }
}

// The stage kernels of the block's form: the fast tiles where they serve
// the block (its fused hydro PLM stage) with binary32 faces in the mixed form
// and binary64 faces in the fast binary64 form; else binary32 fluxes (mixed)
// or the plain binary64 kernels.
void run_stage_kernels(const StageArgs& s, bool mixed)
{
    const bool fast_tile = s.fused && fast64_supported(s.d);
    if (mixed) fast_tile ? stage_kernels_fast32(s) : stage_kernels_f32(s);
    else if (s.d->precision == MARA_PREC_FP64_FAST && fast_tile) stage_kernels_fast64(s);
    else if (s.d->precision == MARA_PREC_FP64_FAST && s.mhd2d) stage_kernels_fast64_mhd2d(s);
    else stage_kernels<double>(s);
}

} // namespace

void mara::stage_round(const StageArgs& s, bool mixed, int* per_cu, int* per_layer)
{
    const bool fast_tile = s.fused && fast64_supported(s.d);
    if (fast_tile && (mixed || s.d->precision == MARA_PREC_FP64_FAST)) stage_round_fast(s, mixed, per_cu, per_layer);
    else if (mixed) stage_round_f32(s, per_cu, per_layer);
    else fused_round_r<double>(*s.d, *s.G, per_cu, per_layer);
}

namespace {

// The buffers stage rk reads and writes.
void stage_io(const mara_block* b, int rk, double** Pin, double** Pout)
{
    const int last = rk == b->d.rk_order - 1;
    *Pin = rk == 0 ? b->Pa : b->Pb;
    *Pout = last ? b->Pa : b->Pb;
    if (b->fused || b->pencil || b->mhd2d)
    {
        // the fused and pencil kernels read neighbouring zones of their input
        // while other work-groups write their output, so input and output
        // never alias: RK2 Pa -> Pb -> Pa, RK3 Pa -> Pb -> Pc -> Pa
        *Pin = rk == 0 ? b->Pa : (rk == 1 ? b->Pb : b->Pc);
        *Pout = last ? b->Pa : (rk == 0 ? b->Pb : b->Pc);
    }
}

// One RK stage: fluxes, CT, update with ghost fill.  With courant_out the last
// stage also reduces the next step's CFL time step into the control bins.
void launch_stage(mara_block* b, int rk, int courant_out)
{
    static const double b1[1] = {1.0};
    static const double b2[2] = {1.0, 0.5};
    static const double b3[3] = {1.0, 1. / 4, 2. / 3};
    const double* bb = b->d.rk_order == 1 ? b1 : b->d.rk_order == 2 ? b2 : b3;
    const int last = rk == b->d.rk_order - 1;
    double* Pin;
    double* Pout;
    stage_io(b, rk, &Pin, &Pout);
    const bool mixed = b->d.precision == MARA_PREC_MIXED;
    const bool wrap = stage_wraps(b);
    // a stage that reads guards starts from filled ones
    if (rk == 0 && !wrap) fill_guards(b);

    // NewtonianMHD cooling (ConservationLaw::addSourceTerms) into S
    const bool cooling = b->G.nq == 9;
    if (cooling)
    {
        const long long zones = (long long)b->G.n[0] * b->G.n[1] * b->G.n[2];
        k_cooling_source<<<(unsigned)((zones + 255) / 256), 256, 0, b->stream>>>(Pin, b->Ssrc, b->G);
    }
    // source terms of the stage's primitives (MeshOperator::generateSourceTerms)
    if (b->src_fn)
    {
        b->src_args.P = Pin;
        b->src_args.S = b->Ssrc;
        void* args[] = {&b->src_args};
        const long long zones = (long long)b->G.n[0] * b->G.n[1] * b->G.n[2];
        (void)hipModuleLaunchKernel(b->src_fn, (unsigned)((zones + 255) / 256), 1, 1, 256, 1, 1, 0, b->stream, args,
                                    nullptr);
    }
    const bool timed = b->timing && b->ev_used < mara_block::kTimingPairs;
    if (timed) hipEventRecord(b->ev[2 * b->ev_used], b->stream);
    StageArgs s;
    s.d = &b->d;
    s.G = &b->G;
    s.stream = b->stream;
    s.Pin = Pin;
    s.P0 = b->Pa;
    s.U = b->U;
    s.Pout = Pout;
    for (int a = 0; a < 3; ++a) s.F[a] = b->F[a];
    s.b = bb[rk];
    s.first = rk == 0;
    s.write_u = !last;
    s.cout = last && courant_out;
    s.ctl = b->ctl;
    s.fused = b->fused;
    s.ct = b->S.ct;
    s.pencil = b->pencil;
    s.mhd2d = b->mhd2d;
    s.kz = b->kz;
    s.rows = b->rows;
    s.S = b->src_fn || cooling ? b->Ssrc : nullptr;
    s.wrap = wrap;
    b->vis_args.P = Pin;
    s.after_flux = b->vis_fn ? launch_viscous : nullptr;
    for (int a = 0; a < 3; ++a) s.Fv[a] = b->vis_fn ? b->Fv[a] : nullptr;
    s.after_flux_ctx = b;
    run_stage_kernels(s, mixed);
    if (timed)
    {
        hipEventRecord(b->ev[2 * b->ev_used + 1], b->stream);
        b->ev_used += 1;
    }
    // the flux-array update writes the ghost zones of Pout itself
    // (write_guard_images); the fused kernels leave them to a ghost-fill pass,
    // and a group member's guards come from the group's exchange after the
    // stage (the images the update wrote are then overwritten).  The PLM
    // tiles of a periodic block read no guards: their output's are filled
    // when something reads them (fill_guards).
    if (b->fused && !b->in_group)
    {
        if (!wrap) launch_bc(b, Pout);
        else if (last) b->guards_stale = true;
    }
}

// ---- tuning table ----------------------------------------------------------
// Shapes chosen by the autotuner, keyed by what selects and sizes a block's
// stage kernels: the kernel law (cooling included), reconstruction, Riemann
// solver, flux-CT, precision, RK order, cell counts and the stage path the
// block takes (the same knob means different tiles on different paths).
std::mutex tune_mutex;
std::map<std::string, mara_shape> tune_table;

enum StagePath { PATH_FLUX_ARRAYS = 0, PATH_FUSED = 1, PATH_PENCIL = 2, PATH_MHD2D = 3 };

// the path a block with no source or viscous hook takes
int desc_path(const mara_desc& d)
{
    if (fused_supported(&d)) return PATH_FUSED;
    if (pencil_supported(&d)) return PATH_PENCIL;
    if (mhd2d_supported(&d)) return PATH_MHD2D;
    return PATH_FLUX_ARRAYS;
}

std::string tune_key(const mara_desc& d, int path)
{
    char k[160];
    std::snprintf(k, sizeof k, "%d %d %d %d %d %d %d %d %d p%d", law_code(d), d.recon, d.riemann,
                  d.law != MARA_LAW_HYDRO && d.field_ct ? 1 : 0, d.precision, d.rk_order, d.cells[0], d.cells[1],
                  d.cells[2], path);
    return k;
}

bool tune_find(const mara_desc& d, mara_shape* s)
{
    std::lock_guard<std::mutex> lock(tune_mutex);
    auto it = tune_table.find(tune_key(d, desc_path(d)));
    if (it == tune_table.end()) return false;
    *s = it->second;
    return true;
}

mara_shape block_shape(const mara_block* b)
{
    mara_shape s;
    s.pencil = b->pencil_shape;
    s.zplanes = b->kz;
    s.rows = b->rows;
    return s;
}

void apply_shape(mara_block* b, const mara_shape& s)
{
    b->pencil_shape = s.pencil;
    if (b->pencil) b->pencil = pencil_length(b->d.law, s.pencil);
    b->kz = s.zplanes;
    b->rows = s.rows;
    // the captured steps hold the old launch shape
    if (b->steps_exec) (void)hipGraphExecDestroy(b->steps_exec);
    b->steps_exec = nullptr;
}

// n device-driven steps: dt from the CFL bins and the clock, the RK stages
// (the last one reduces the next step's CFL), then the clock advance, with
// each step's end and the next one's start in one launch (k_ctl_step_next).
void enqueue_steps(mara_block* b, int n)
{
    if (n <= 0) return;
    k_ctl_step_start<<<1, 64, 0, b->stream>>>(b->ctl);
    for (int s = 0; s < n; ++s)
    {
        for (int rk = 0; rk < b->d.rk_order; ++rk) launch_stage(b, rk, 1);
        if (s + 1 < n) k_ctl_step_next<<<1, 64, 0, b->stream>>>(b->ctl);
        else k_ctl_step_end<<<1, 1, 0, b->stream>>>(b->ctl);
    }
}

// Capture kGraphSteps steps of the block's stream into an executable graph:
// the launches of enqueue_steps.  Nothing in a step is a memset or a host
// transfer.
int capture_steps(mara_block* b)
{
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    enqueue_steps(b, mara_block::kGraphSteps);
    const hipError_t launched = hipGetLastError();
    const hipError_t ended = hipStreamEndCapture(b->stream, &g);
    if (!hip_ok(launched) || !hip_ok(ended))
    {
        if (g) (void)hipGraphDestroy(g);
        return MARA_ERR_HIP;
    }
    const hipError_t made = hipGraphInstantiate(&b->steps_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (!hip_ok(made))
    {
        b->steps_exec = nullptr;
        return MARA_ERR_HIP;
    }
    return MARA_OK;
}

int read_ctl(mara_block* b)
{
    HIP_TRY(hipMemcpyAsync(b->ctl_host, b->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return MARA_OK;
}

int fill_report(mara_block* b, mara_report* r) { return host::fill_report(*b->ctl_host, b->G, r); }

} // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* mara_version(void) { return "mara2-amd 0.1 (gfx950)"; }

const char* mara_error_string(int status)
{
    switch (status)
    {
        case MARA_OK: return "ok";
        case MARA_ERR_INVALID_ARGUMENT: return "invalid argument";
        case MARA_ERR_GUARD_TOO_SMALL: return "Boundary region of mesh data is smaller than the scheme's stencil";
        case MARA_ERR_UNSUPPORTED: return "configuration not supported";
        case MARA_ERR_HIP: return "HIP runtime error";
        case MARA_ERR_INVERSION_FAILURE: return "conserved to primitive inversion failed";
        case MARA_ERR_SUPERLUMINAL: return "got superluminal velocity";
        case MARA_ERR_BAD_EIGENVALUES: return "got bad eigenvalues";
        case MARA_ERR_NO_DEVICE: return "no usable GPU device";
        case MARA_ERR_CAPACITY: return "request exceeds capacity";
        default: return "unknown status";
    }
}

int mara_num_fields(int law) { return law_nq(law); }

int mara_desc_num_fields(const mara_desc* d) { return d ? law_nq(law_code(*d)) : -1; }
int mara_stencil_size(int recon) { return recon_ng(recon); }

long long mara_desc_device_bytes(const mara_desc* d)
{
    if (validate(d) != MARA_OK) return -1;
    return device_bytes(d, sizes_of(d));
}

int mara_block_create(const mara_desc* d, mara_block** out)
{
    if (!out) return MARA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int v = validate(d);
    if (v) return v;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MARA_ERR_NO_DEVICE;

    mara_block* b = new (std::nothrow) mara_block();
    if (!b) return MARA_ERR_HIP;
    b->d = *d;
    b->S = sizes_of(d);
    b->ng = recon_ng(d->recon);
    b->fused = fused_supported(d);
    b->pencil = pencil_supported(d) ? PENCIL_DEFAULT : 0;
    b->mhd2d = mhd2d_supported(d);
    if (b->mhd2d) b->rows = MHD_ROWS_DEFAULT;
    b->G = make_geom(*d);
    {
        mara_shape tuned;
        if (tune_find(*d, &tuned)) apply_shape(b, tuned);
    }

    auto fail = [&](int code) { mara_block_destroy(b); return code; };
    if (!hip_ok(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking))) return fail(MARA_ERR_HIP);
    const size_t cellbytes = (size_t)b->S.nq * b->S.fs * sizeof(double);
    if (!hip_ok(hipMalloc(&b->Pa, cellbytes))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipMalloc(&b->Pb, cellbytes))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipMalloc(&b->U, cellbytes))) return fail(MARA_ERR_HIP);
    if ((b->fused || b->pencil || b->mhd2d) && d->rk_order == 3)
    {
        if (!hip_ok(hipMalloc(&b->Pc, cellbytes))) return fail(MARA_ERR_HIP);
        if (!hip_ok(hipMemset(b->Pc, 0, cellbytes))) return fail(MARA_ERR_HIP);
    }
    if (!b->fused && !b->pencil && !b->mhd2d)
    {
        for (int a = 0; a < 3; ++a)
        {
            if (!b->S.fleshed[a]) continue;
            const size_t fb = (size_t)b->S.nq * b->S.ffs * b->S.real;
            if (!hip_ok(hipMalloc(&b->F[a], fb))) return fail(MARA_ERR_HIP);
        }
    }
    if (b->G.nq == 9)
    {
        // the cooling terms of each stage (k_cooling_source)
        if (!hip_ok(hipMalloc(&b->Ssrc, cellbytes))) return fail(MARA_ERR_HIP);
        if (!hip_ok(hipMemset(b->Ssrc, 0, cellbytes))) return fail(MARA_ERR_HIP);
    }
    if (d->geometry == MARA_GEOM_SPHERICAL)
    {
        const std::vector<double> tab = spherical_tables(*d, b->G);
        const size_t tb = tab.size() * sizeof(double);
        if (!hip_ok(hipMalloc(&b->sph, tb))) return fail(MARA_ERR_HIP);
        if (!hip_ok(hipMemcpy(b->sph, tab.data(), tb, hipMemcpyHostToDevice))) return fail(MARA_ERR_HIP);
        b->G.sph = b->sph;
    }
    if (!hip_ok(hipMalloc(&b->ctl, sizeof(Ctl)))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipMalloc(&b->cbins, CBINS * CBIN_STRIDE * sizeof(unsigned long long)))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipHostMalloc((void**)&b->ctl_host, sizeof(Ctl), hipHostMallocDefault))) return fail(MARA_ERR_HIP);
    std::memset(b->ctl_host, 0, sizeof(Ctl));
    b->ctl_host->active = 1;
    b->ctl_host->first_fail = ~0ull;
    b->ctl_host->cbins = b->cbins;
    if (!hip_ok(hipMemcpy(b->ctl, b->ctl_host, sizeof(Ctl), hipMemcpyHostToDevice))) return fail(MARA_ERR_HIP);
    // guard zones and never-written faces start at zero
    if (!hip_ok(hipMemset(b->Pa, 0, cellbytes))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipMemset(b->Pb, 0, cellbytes))) return fail(MARA_ERR_HIP);
    b->bytes = device_bytes(d, b->S);
    *out = b;
    return MARA_OK;
}

int mara_block_destroy(mara_block* b)
{
    if (!b) return MARA_ERR_INVALID_ARGUMENT;
    if (b->stream) hipStreamSynchronize(b->stream);
    hipFree(b->Pa);
    hipFree(b->Pb);
    hipFree(b->Pc);
    hipFree(b->U);
    for (int a = 0; a < 3; ++a) hipFree(b->F[a]);
    for (int a = 0; a < 3; ++a) hipFree(b->Fv[a]);
    hipFree(b->ctl);
    hipFree(b->cbins);
    if (b->steps_exec) (void)hipGraphExecDestroy(b->steps_exec);
    if (b->src_module) (void)hipModuleUnload(b->src_module);
    if (b->vis_module) (void)hipModuleUnload(b->vis_module);
    hipFree(b->Ssrc);
    hipFree(b->sph);
    if (b->ctl_host) hipHostFree(b->ctl_host);
    for (int e = 0; e < 2 * mara_block::kTimingPairs; ++e)
        if (b->ev[e]) hipEventDestroy(b->ev[e]);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return MARA_OK;
}

long long mara_block_device_bytes(const mara_block* b) { return b ? b->bytes : -1; }

int mara_block_get_shape(const mara_block* b, mara_shape* shape)
{
    if (!b || !shape) return MARA_ERR_INVALID_ARGUMENT;
    *shape = block_shape(b);
    return MARA_OK;
}

int mara_block_set_shape(mara_block* b, const mara_shape* shape)
{
    if (!b || !shape || !valid_shape(*shape)) return MARA_ERR_INVALID_ARGUMENT;
    apply_shape(b, *shape);
    return MARA_OK;
}

int mara_block_autotune(mara_block* b, int repeats, mara_shape* chosen, double* best_ms)
{
    if (!b || repeats < 1) return MARA_ERR_INVALID_ARGUMENT;
    // the candidates of the one shape the block's stages use
    std::vector<mara_shape> cands;
    const mara_shape cur = block_shape(b);
    if (b->pencil && b->d.law != MARA_LAW_RMHD)
        for (int p : PENCIL_SHAPES) { mara_shape s = cur; s.pencil = p; cands.push_back(s); }
    else if (b->fused && b->G.N[2] > 1)
    {
        for (int z : KZ_SHAPES) { mara_shape s = cur; s.zplanes = z; cands.push_back(s); }
        // z-plane counts whose tile count is a whole number (1 to 3) of device
        // rounds, each round every CU holding as many work-groups as the
        // kernel's occupancy allows: no partly filled last round, and fewer
        // tiles re-reading the z window at their start
        StageArgs q;
        q.d = &b->d;
        q.G = &b->G;
        int per_cu = 0, per_layer = 0, cus = 0, dev = 0;
        stage_round(q, b->d.precision == MARA_PREC_MIXED, &per_cu, &per_layer);
        if (hipGetDevice(&dev) != hipSuccess
            || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 0;
        (void)hipGetLastError();
        for (int r = 1; r <= 3 && per_cu > 0 && per_layer > 0 && cus > 0; ++r)
        {
            const long long layers = (long long)r * per_cu * cus / per_layer;
            if (layers < 1) continue;
            const int z = (int)((b->G.n[2] + layers - 1) / layers);
            bool have = z < 4 || z > KZ_MAX;
            for (const mara_shape& c : cands) have = have || c.zplanes == z;
            if (!have) { mara_shape s = cur; s.zplanes = z; cands.push_back(s); }
        }
    }
    else if (!b->fused && !b->pencil)  // flux-array rows or fused 2-D MHD tile height
        for (int r : ROWS_SHAPES) { mara_shape s = cur; s.rows = r; cands.push_back(s); }
    else
        cands.push_back(cur);

    // scratch control block: active, dt = 0 (the stage reproduces U0), no CFL
    Ctl h;
    std::memset(&h, 0, sizeof h);
    h.active = 1;
    h.t_final = 1.0;
    h.max_steps = 1;
    h.first_fail = ~0ull;
    h.cbins = b->cbins;
    Ctl* sc = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int status = MARA_OK;
    double best = 1e300;
    mara_shape pick = cur;
    auto cleanup = [&]() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (sc) (void)hipFree(sc);
    };
    if (!hip_ok(hipMalloc(&sc, sizeof(Ctl))) || !hip_ok(hipMemcpy(sc, &h, sizeof(Ctl), hipMemcpyHostToDevice))
        || !hip_ok(hipEventCreate(&e0)) || !hip_ok(hipEventCreate(&e1)))
    {
        cleanup();
        return MARA_ERR_HIP;
    }
    // candidates in turn, three rounds, each candidate's fastest round counts:
    // a slow moment of the device then penalises one round, not one candidate
    std::vector<double> fastest(cands.size(), 1e300);
    const size_t rounds = cands.size() > 1 ? 3 : 1;
    for (size_t it = 0; it < rounds * cands.size(); ++it)
    {
        const size_t ci = it % cands.size();
        const mara_shape& c = cands[ci];
        StageArgs s;
        s.d = &b->d;
        s.G = &b->G;
        s.stream = b->stream;
        s.Pin = b->Pa;  // stage 1 reads the state and writes only scratch
        s.P0 = b->Pa;
        s.U = b->U;
        s.Pout = b->Pb;
        for (int a = 0; a < 3; ++a) s.F[a] = b->F[a];
        s.b = 1.0;
        s.first = 1;
        s.write_u = 1;
        s.cout = 0;
        s.ctl = sc;
        s.fused = b->fused;
        s.ct = b->S.ct;
        s.pencil = b->pencil ? pencil_length(b->d.law, c.pencil) : 0;
        s.mhd2d = b->mhd2d;
        s.kz = c.zplanes;
        s.rows = c.rows;
        s.S = b->src_fn || b->G.nq == 9 ? b->Ssrc : nullptr;
        s.wrap = stage_wraps(b);
        s.after_flux = nullptr;  // the tuner times the stage kernels only
        s.after_flux_ctx = nullptr;
        const bool mixed = b->d.precision == MARA_PREC_MIXED;
        for (int r = 0; r <= repeats; ++r)
        {
            if (r == 1) (void)hipEventRecord(e0, b->stream);
            run_stage_kernels(s, mixed);
        }
        (void)hipEventRecord(e1, b->stream);
        if (!hip_ok(hipGetLastError()) || !hip_ok(hipEventSynchronize(e1)))
        {
            status = MARA_ERR_HIP;
            break;
        }
        float ms = 0.0f;
        if (!hip_ok(hipEventElapsedTime(&ms, e0, e1)))
        {
            status = MARA_ERR_HIP;
            break;
        }
        fastest[ci] = std::min(fastest[ci], (double)ms / repeats);
    }
    cleanup();
    if (status) return status;
    for (size_t ci = 0; ci < cands.size(); ++ci)
        if (fastest[ci] < best)
        {
            best = fastest[ci];
            pick = cands[ci];
        }
    apply_shape(b, pick);
    {
        std::lock_guard<std::mutex> lock(tune_mutex);
        // the path the block took while it was timed (a hook moves it to the flux arrays)
        const int path = b->fused ? PATH_FUSED : b->pencil ? PATH_PENCIL : b->mhd2d ? PATH_MHD2D : PATH_FLUX_ARRAYS;
        tune_table[tune_key(b->d, path)] = pick;
    }
    if (chosen) *chosen = pick;
    if (best_ms) *best_ms = best;
    return MARA_OK;
}

int mara_tune_lookup(const mara_desc* d, mara_shape* shape)
{
    if (!d || !shape) return MARA_ERR_INVALID_ARGUMENT;
    return tune_find(*d, shape) ? 1 : 0;
}

int mara_tune_save(const char* path)
{
    if (!path) return MARA_ERR_INVALID_ARGUMENT;
    std::FILE* f = std::fopen(path, "w");
    if (!f) return MARA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(tune_mutex);
    std::fprintf(f, "# kernel-law recon riemann ct precision rk_order n0 n1 n2 p<path> : pencil zplanes rows\n"
                    "# path: 0 flux arrays, 1 fused hydro, 2 1-D pencils, 3 fused 2-D MHD\n");
    for (auto& e : tune_table)
        std::fprintf(f, "%s : %d %d %d\n", e.first.c_str(), e.second.pencil, e.second.zplanes, e.second.rows);
    std::fclose(f);
    return MARA_OK;
}

int mara_tune_load(const char* path)
{
    if (!path) return MARA_ERR_INVALID_ARGUMENT;
    std::FILE* f = std::fopen(path, "r");
    if (!f) return MARA_ERR_INVALID_ARGUMENT;
    char line[256];
    int status = MARA_OK;
    std::map<std::string, mara_shape> got;
    while (std::fgets(line, sizeof line, f))
    {
        if (line[0] == '#' || line[0] == '\n') continue;
        int k[10];
        mara_shape s;
        if (std::sscanf(line, "%d %d %d %d %d %d %d %d %d p%d : %d %d %d", &k[0], &k[1], &k[2], &k[3], &k[4], &k[5],
                        &k[6], &k[7], &k[8], &k[9], &s.pencil, &s.zplanes, &s.rows) != 13
            || !valid_shape(s) || k[9] < PATH_FLUX_ARRAYS || k[9] > PATH_MHD2D)
        {
            status = MARA_ERR_INVALID_ARGUMENT;
            break;
        }
        char key[160];
        std::snprintf(key, sizeof key, "%d %d %d %d %d %d %d %d %d p%d", k[0], k[1], k[2], k[3], k[4], k[5], k[6],
                      k[7], k[8], k[9]);
        got[key] = s;
    }
    std::fclose(f);
    if (status) return status;
    std::lock_guard<std::mutex> lock(tune_mutex);
    for (auto& e : got) tune_table[e.first] = e.second;
    return MARA_OK;
}

int mara_block_upload_prim(mara_block* b, const double* P)
{
    if (!b || !P) return MARA_ERR_INVALID_ARGUMENT;
    const size_t n = (size_t)b->S.ncell * b->S.nq;
    HIP_TRY(hipMemcpyAsync(b->U, P, n * sizeof(double), hipMemcpyHostToDevice, b->stream));
    k_aos_to_soa<<<grid_for(b->S.ncell), kThreads, 0, b->stream>>>(b->U, b->Pa, b->G);
    b->guards_stale = false;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(b->stream));
    return MARA_OK;
}

int mara_block_download_prim(mara_block* b, double* P)
{
    if (!b || !P) return MARA_ERR_INVALID_ARGUMENT;
    const size_t n = (size_t)b->S.ncell * b->S.nq;
    fill_guards(b);
    k_soa_to_aos<<<grid_for(b->S.ncell), kThreads, 0, b->stream>>>(b->Pa, b->U, b->G);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(P, b->U, n * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return MARA_OK;
}

int mara_block_apply_bc(mara_block* b)
{
    if (!b) return MARA_ERR_INVALID_ARGUMENT;
    k_ctl_begin_courant<<<1, 1, 0, b->stream>>>(b->ctl);  // sets active
    launch_bc(b, b->Pa);
    b->guards_stale = false;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(b->stream));
    return MARA_OK;
}

int mara_block_advance(mara_block* b, double t0, double dt, mara_report* report)
{
    (void)t0;
    if (!b) return MARA_ERR_INVALID_ARGUMENT;
    k_ctl_begin_advance<<<1, 1, 0, b->stream>>>(b->ctl, dt);
    for (int rk = 0; rk < b->d.rk_order; ++rk) launch_stage(b, rk, 0);
    HIP_TRY(hipGetLastError());
    int e = read_ctl(b);
    if (e) return e;
    return fill_report(b, report);
}

int mara_block_courant_dt(mara_block* b, double* dt)
{
    if (!b || !dt) return MARA_ERR_INVALID_ARGUMENT;
    k_ctl_begin_courant<<<1, 1, 0, b->stream>>>(b->ctl);
    launch_courant(b);
    k_ctl_fold_bins<<<1, 1, 0, b->stream>>>(b->ctl);
    HIP_TRY(hipGetLastError());
    int e = read_ctl(b);
    if (e) return e;
    union { unsigned long long u; double d; } cv;
    cv.u = b->ctl_host->courant_bits;
    *dt = cv.d;
    if (b->ctl_host->superluminal) return MARA_ERR_SUPERLUMINAL;
    if (b->ctl_host->eigen_fail) return MARA_ERR_BAD_EIGENVALUES;
    return MARA_OK;
}

int mara_block_monopole_max(mara_block* b, double* max_abs)
{
    if (!b || !max_abs) return MARA_ERR_INVALID_ARGUMENT;
    if (b->d.law == MARA_LAW_HYDRO) return MARA_ERR_UNSUPPORTED;
    k_ctl_begin_monopole<<<1, 1, 0, b->stream>>>(b->ctl);
    const long long total = (long long)b->G.n[0] * b->G.n[1] * b->G.n[2];
    k_monopole<<<grid_for(total), kThreads, 0, b->stream>>>(b->Pa, b->G, b->ctl);
    HIP_TRY(hipGetLastError());
    int e = read_ctl(b);
    if (e) return e;
    union { unsigned long long u; double d; } cv;
    cv.u = b->ctl_host->monopole_bits;
    *max_abs = cv.d;
    return MARA_OK;
}

int mara_block_run(mara_block* b, double cfl, double t_final, int max_steps,
                   double* t, int* steps, mara_report* report)
{
    if (!b || !t || !steps || max_steps < 0) return MARA_ERR_INVALID_ARGUMENT;
    // stage timing records events around each stage launch, which a graph
    // replay would overwrite: timed runs launch eagerly
    const bool graph = b->graphs && !b->timing;
    if (graph && !b->steps_exec)
    {
        const int e = capture_steps(b);
        if (e) return e;
    }
    k_ctl_begin_run<<<1, 1, 0, b->stream>>>(b->ctl, *t, t_final, cfl, max_steps);
    // the first step's CFL; later steps take it from the previous step's last stage
    launch_courant(b);
    constexpr int kChunk = 64;
    int queued = 0, chunk = kChunk;
    while (queued < max_steps)
    {
        // enqueue a chunk of steps; steps after the stop condition are no-ops
        if (chunk > max_steps - queued) chunk = max_steps - queued;
        int s = 0;
        if (graph)
            for (; s + mara_block::kGraphSteps <= chunk; s += mara_block::kGraphSteps)
                HIP_TRY(hipGraphLaunch(b->steps_exec, b->stream));
        enqueue_steps(b, chunk - s);
        queued += chunk;
        HIP_TRY(hipGetLastError());
        int e = read_ctl(b);
        if (e) return e;
        const Ctl& c = *b->ctl_host;
        if (!c.active || !(c.t < c.t_final) || c.steps >= max_steps) break;
        // size the next chunk to the steps the clock still needs at the last
        // step's dt (plus a margin), so few no-op steps are queued at the end
        const double left = (c.t_final - c.t) / c.dt;
        chunk = kChunk;
        if (c.dt > 0 && left + 8 < kChunk) chunk = (int)left + 8;
    }
    int e = read_ctl(b);
    if (e) return e;
    *t = b->ctl_host->t;
    *steps = b->ctl_host->steps;
    return fill_report(b, report);
}

int mara_block_set_source(mara_block* b, const char* code_object, const char* kernel, const double* params,
                          int n_params)
{
    if (!b || n_params < 0 || n_params > MARA_SOURCE_MAX_PARAMS || (n_params > 0 && !params)
        || (code_object && !kernel))
        return MARA_ERR_INVALID_ARGUMENT;
    // the cooling terms use the S array; the reference adds them to L before
    // the source function's, an order one array cannot keep
    if (code_object && b->G.nq == 9) return MARA_ERR_UNSUPPORTED;
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->src_module) (void)hipModuleUnload(b->src_module);
    b->src_module = nullptr;
    b->src_fn = nullptr;
    if (b->steps_exec) (void)hipGraphExecDestroy(b->steps_exec);  // captured without / with the old source
    b->steps_exec = nullptr;
    if (!code_object)
    {
        b->fused = fused_supported(&b->d) && !b->vis_fn;
        return MARA_OK;
    }
    if (hipModuleLoad(&b->src_module, code_object) != hipSuccess)
    {
        std::fprintf(stderr, "mara_gpu: cannot load source code object %s\n", code_object);
        (void)hipGetLastError();  // the failure is reported here, not by the next launch check
        b->src_module = nullptr;
        return MARA_ERR_INVALID_ARGUMENT;
    }
    if (hipModuleGetFunction(&b->src_fn, b->src_module, kernel) != hipSuccess)
    {
        std::fprintf(stderr, "mara_gpu: no kernel %s in %s\n", kernel, code_object);
        (void)hipGetLastError();
        (void)hipModuleUnload(b->src_module);
        b->src_module = nullptr;
        b->src_fn = nullptr;
        return MARA_ERR_INVALID_ARGUMENT;
    }
    const size_t cellbytes = (size_t)b->S.nq * b->S.fs * sizeof(double);
    if (!b->Ssrc)
    {
        HIP_TRY(hipMalloc(&b->Ssrc, cellbytes));
        HIP_TRY(hipMemset(b->Ssrc, 0, cellbytes));
        b->bytes += (long long)cellbytes;
    }
    // the fused hydro kernels update zones without a source term: such a
    // block takes the flux-array path while a source is set
    if (b->fused)
    {
        for (int a = 0; a < 3; ++a)
        {
            if (!b->S.fleshed[a] || b->F[a]) continue;
            const size_t fb = (size_t)b->S.nq * b->S.ffs * b->S.real;
            HIP_TRY(hipMalloc(&b->F[a], fb));
            b->bytes += (long long)fb;
        }
        b->fused = false;
    }
    mara_source_args& a = b->src_args;
    a.fs = b->G.fs;
    for (int e = 0; e < 3; ++e)
    {
        a.N[e] = b->G.N[e];
        a.g[e] = b->G.g[e];
        a.n[e] = b->G.n[e];
        a.lower[e] = b->d.lower[e];
        a.dx[e] = (b->d.upper[e] - b->d.lower[e]) / b->d.cells[e];  // CartesianMeshGeometry
        a.centre[e] = nullptr;
    }
    if (b->sph)
    {
        // SphericalMeshGeometry::coordinateAtIndex of the interior cells
        const SphView s(b->G);
        a.centre[0] = s.rc;
        a.centre[1] = s.qc;
        a.centre[2] = s.pc;
    }
    a.nq = b->S.nq;
    for (int q = 0; q < MARA_SOURCE_MAX_PARAMS; ++q) a.params[q] = q < n_params ? params[q] : 0.0;
    return MARA_OK;
}

int mara_block_set_source_aux(mara_block* b, const double* t)
{
    if (!b || !t) return MARA_ERR_INVALID_ARGUMENT;
    for (int q = 0; q < 8; ++q) b->src_args.t[q] = b->vis_args.t[q] = t[q];
    if (b->steps_exec) (void)hipGraphExecDestroy(b->steps_exec);  // launch records are captured by value
    b->steps_exec = nullptr;
    return MARA_OK;
}

int mara_block_set_viscous_flux(mara_block* b, const char* code_object, const char* kernel, const double* params,
                                int n_params)
{
    if (!b || n_params < 0 || n_params > MARA_SOURCE_MAX_PARAMS || (n_params > 0 && !params)
        || (code_object && !kernel))
        return MARA_ERR_INVALID_ARGUMENT;
    if (code_object && b->sph) return MARA_ERR_UNSUPPORTED;
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->vis_module) (void)hipModuleUnload(b->vis_module);
    b->vis_module = nullptr;
    b->vis_fn = nullptr;
    if (b->steps_exec) (void)hipGraphExecDestroy(b->steps_exec);  // captured without / with the old hook
    b->steps_exec = nullptr;
    if (!code_object)
    {
        // back to the block's own stage kernels
        b->fused = fused_supported(&b->d) && !b->src_fn;
        b->pencil = pencil_supported(&b->d) ? pencil_length(b->d.law, b->pencil_shape) : 0;
        b->mhd2d = mhd2d_supported(&b->d);
        return MARA_OK;
    }
    if (hipModuleLoad(&b->vis_module, code_object) != hipSuccess)
    {
        std::fprintf(stderr, "mara_gpu: cannot load viscous-flux code object %s\n", code_object);
        (void)hipGetLastError();  // the failure is reported here, not by the next launch check
        b->vis_module = nullptr;
        return MARA_ERR_INVALID_ARGUMENT;
    }
    if (hipModuleGetFunction(&b->vis_fn, b->vis_module, kernel) != hipSuccess)
    {
        std::fprintf(stderr, "mara_gpu: no kernel %s in %s\n", kernel, code_object);
        (void)hipGetLastError();
        (void)hipModuleUnload(b->vis_module);
        b->vis_module = nullptr;
        b->vis_fn = nullptr;
        return MARA_ERR_INVALID_ARGUMENT;
    }
    // the flux-array path: flux arrays of every axis with zones, and a
    // binary64 staging copy where the arrays are binary32 or flux-CT must be
    // applied before the hook
    const bool staged = b->d.precision == MARA_PREC_MIXED || b->S.ct;
    for (int a = 0; a < 3; ++a)
    {
        if (!b->S.fleshed[a]) continue;
        if (!b->F[a])
        {
            const size_t fb = (size_t)b->S.nq * b->S.ffs * b->S.real;
            HIP_TRY(hipMalloc(&b->F[a], fb));
            b->bytes += (long long)fb;
        }
        if (staged && !b->Fv[a])
        {
            const size_t fb = (size_t)b->S.nq * b->S.ffs * sizeof(double);
            HIP_TRY(hipMalloc(&b->Fv[a], fb));
            b->bytes += (long long)fb;
        }
    }
    b->fused = false;
    b->pencil = 0;
    b->mhd2d = false;
    mara_flux_args& a = b->vis_args;
    a.fs = b->G.fs;
    a.ffs = b->G.ffs;
    a.nq = b->S.nq;
    b->vis_faces = 0;
    for (int e = 0; e < 3; ++e)
    {
        a.F[e] = !b->S.fleshed[e] ? nullptr : static_cast<double*>(staged ? b->Fv[e] : b->F[e]);
        a.N[e] = b->G.N[e];
        a.g[e] = b->G.g[e];
        a.n[e] = b->G.n[e];
        a.lower[e] = b->d.lower[e];
        a.dx[e] = (b->d.upper[e] - b->d.lower[e]) / b->d.cells[e];  // CartesianMeshGeometry
        if (a.F[e])
            b->vis_faces += (long long)(b->G.n[0] + (e == 0)) * (b->G.n[1] + (e == 1)) * (b->G.n[2] + (e == 2));
    }
    for (int q = 0; q < 8; ++q) a.t[q] = b->src_args.t[q];
    for (int q = 0; q < MARA_SOURCE_MAX_PARAMS; ++q) a.params[q] = q < n_params ? params[q] : 0.0;
    return MARA_OK;
}

int mara_block_set_graphs(mara_block* b, int enable)
{
    if (!b) return MARA_ERR_INVALID_ARGUMENT;
    b->graphs = enable != 0;
    return MARA_OK;
}

int mara_block_set_timing(mara_block* b, int enable)
{
    if (!b) return MARA_ERR_INVALID_ARGUMENT;
    if (enable && !b->ev[0])
        for (int e = 0; e < 2 * mara_block::kTimingPairs; ++e) HIP_TRY(hipEventCreate(&b->ev[e]));
    b->timing = enable != 0;
    b->ev_used = 0;
    b->timed_ms = 0.0;
    b->timed_launches = 0;
    return MARA_OK;
}

int mara_block_stage_kernel_time(mara_block* b, double* avg_ms, long long* launches, const char** name)
{
    if (!b) return MARA_ERR_INVALID_ARGUMENT;
    if (b->ev_used > 0)
    {
        HIP_TRY(hipEventSynchronize(b->ev[2 * b->ev_used - 1]));
        for (int p = 0; p < b->ev_used; ++p)
        {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, b->ev[2 * p], b->ev[2 * p + 1]));
            b->timed_ms += ms;
            b->timed_launches += 1;
        }
        b->ev_used = 0;
    }
    if (avg_ms) *avg_ms = b->timed_launches ? b->timed_ms / b->timed_launches : 0.0;
    if (launches) *launches = b->timed_launches;
    if (name)
        *name = b->fused   ? fused_kernel_name(b->d)
                : b->mhd2d ? "k_mhd2d (2-D MHD stage)"
                : b->pencil ? "k_pencil (1-D stage)"
                            : "k_flux+k_update (flux-array stage)";
    return MARA_OK;
}

} // extern "C"

// ===========================================================================
// One step with no device allocation (mara_advance_noalloc)
// ===========================================================================
// Every working buffer of the step lives in this file-scope array, one copy
// per device, held from the code object's load until the process ends.  In
// host code the name `scratch` is a host-side stand-in, not a device
// address: the address of a device's copy comes from hipGetSymbolAddress,
// taken once per device with that device current, and only that address is
// passed to kernels.
static __device__ __align__(256) unsigned char scratch[MARA_SCRATCH_BYTES];

namespace {

constexpr int kScratchDevices = 64;
std::mutex scratch_mutex;
unsigned char* scratch_ptr[kScratchDevices] = {};

// the current device's copy of the scratch array
int scratch_on_device(int dev, unsigned char** out)
{
    if (dev < 0 || dev >= kScratchDevices) return MARA_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev])
    {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
        {
            (void)hipGetLastError();
            return MARA_ERR_HIP;
        }
        scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = scratch_ptr[dev];
    return MARA_OK;
}

// Byte offsets of a step's buffers in the scratch array, each 256-aligned:
// the same buffers mara_block_create allocates for the descriptor.
struct ScratchPlan
{
    long long Pa, Pb, Pc, U, F[3], S, ctl, cbins, total;
};

ScratchPlan plan_scratch(const mara_desc& d, const Sizes& s, bool in_pass)
{
    ScratchPlan p;
    long long at = 0;
    auto take = [&](long long bytes) {
        const long long here = at;
        at += round_up(bytes, 256);
        return here;
    };
    const long long cellbytes = (long long)s.nq * s.fs * 8;
    p.Pa = take(cellbytes);
    p.Pb = take(cellbytes);
    p.U = take(cellbytes);
    p.Pc = in_pass && d.rk_order == 3 ? take(cellbytes) : -1;
    for (int a = 0; a < 3; ++a) p.F[a] = !in_pass && s.fleshed[a] ? take((long long)s.nq * s.ffs * s.real) : -1;
    p.S = s.nq == 9 ? take(cellbytes) : -1;
    p.ctl = take(sizeof(Ctl));
    p.cbins = take((long long)CBINS * CBIN_STRIDE * 8);
    p.total = at;
    return p;
}

} // namespace

// the control block of a fresh step: counters at zero, the CFL bins reset
static __global__ void k_scratch_ctl_reset(Ctl* ctl, unsigned long long* cbins)
{
    const int t = threadIdx.x;
    if (t == 0)
    {
        ctl->dt = 0.0;
        ctl->t = 0.0;
        ctl->t_final = 0.0;
        ctl->cfl = 0.0;
        ctl->courant_bits = DBL_MAX_BITS;
        ctl->monopole_bits = 0;
        ctl->first_fail = ~0ull;
        ctl->c2p_fail = 0;
        ctl->floor_applied = 0;
        ctl->eigen_fail = 0;
        ctl->superluminal = 0;
        ctl->active = 1;
        ctl->steps = 0;
        ctl->max_steps = 0;
        ctl->pad = 0;
        ctl->cbins = cbins;
    }
    for (int b = t; b < CBINS; b += blockDim.x) cbins[b * CBIN_STRIDE] = DBL_MAX_BITS;
}

static __global__ void __launch_bounds__(256) k_scratch_zero(double* __restrict__ p, long long n)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        p[t] = 0.0;
}

extern "C" int mara_advance_noalloc(const mara_desc* d, const double* P_in, double* P_out, double t0, double dt,
                                    void* stream, mara_report* report)
{
    (void)t0;
    if (!P_in || !P_out) return MARA_ERR_INVALID_ARGUMENT;
    const int v = validate(d);
    if (v) return v;
    // a spherical block's geometry tables would need a scratch range of their own
    if (d->geometry != MARA_GEOM_CARTESIAN) return MARA_ERR_UNSUPPORTED;
    // the block's state, set up as mara_block_create does, on the stack; its
    // buffers are carved from the scratch array below
    mara_block b;
    b.d = *d;
    b.S = sizes_of(d);
    b.ng = recon_ng(d->recon);
    b.fused = fused_supported(d);
    b.pencil = pencil_supported(d) ? PENCIL_DEFAULT : 0;
    b.mhd2d = mhd2d_supported(d);
    if (b.mhd2d) b.rows = MHD_ROWS_DEFAULT;
    b.G = make_geom(*d);
    const ScratchPlan plan = plan_scratch(*d, b.S, b.fused || b.pencil || b.mhd2d);
    if (plan.total > MARA_SCRATCH_BYTES) return MARA_ERR_CAPACITY;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess)
    {
        (void)hipGetLastError();
        return MARA_ERR_NO_DEVICE;
    }
    unsigned char* base = nullptr;
    const int sa = scratch_on_device(dev, &base);
    if (sa) return sa;
    auto at = [&](long long off) { return off < 0 ? nullptr : reinterpret_cast<double*>(base + off); };
    b.Pa = at(plan.Pa);
    b.Pb = at(plan.Pb);
    b.Pc = at(plan.Pc);
    b.U = at(plan.U);
    for (int a = 0; a < 3; ++a) b.F[a] = plan.F[a] < 0 ? nullptr : static_cast<void*>(base + plan.F[a]);
    b.Ssrc = at(plan.S);
    b.ctl = reinterpret_cast<Ctl*>(base + plan.ctl);
    b.cbins = reinterpret_cast<unsigned long long*>(base + plan.cbins);
    Ctl ctl_host;
    b.ctl_host = &ctl_host;
    b.stream = static_cast<hipStream_t>(stream);
    b.graphs = false;

    // what mara_block_create zeroes, zeroed again: the array keeps the last
    // call's contents
    const long long cellwords = (long long)b.S.nq * b.S.fs;
    const unsigned zgrid = (unsigned)std::min<long long>((cellwords + 255) / 256, 4096);
    k_scratch_ctl_reset<<<1, 64, 0, b.stream>>>(b.ctl, b.cbins);
    k_scratch_zero<<<zgrid, 256, 0, b.stream>>>(b.Pb, cellwords);
    if (b.Pc) k_scratch_zero<<<zgrid, 256, 0, b.stream>>>(b.Pc, cellwords);
    if (b.Ssrc) k_scratch_zero<<<zgrid, 256, 0, b.stream>>>(b.Ssrc, cellwords);
    // MeshData::P -> the device layout, one step, and back
    k_aos_to_soa<<<grid_for(b.S.ncell), kThreads, 0, b.stream>>>(P_in, b.Pa, b.G);
    k_ctl_begin_advance<<<1, 1, 0, b.stream>>>(b.ctl, dt);
    for (int rk = 0; rk < b.d.rk_order; ++rk) launch_stage(&b, rk, 0);
    fill_guards(&b);
    k_soa_to_aos<<<grid_for(b.S.ncell), kThreads, 0, b.stream>>>(b.Pa, P_out, b.G);
    int status = MARA_OK;
    if (!hip_ok(hipGetLastError())) status = MARA_ERR_HIP;
    else if (!hip_ok(hipMemcpyAsync(&ctl_host, b.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, b.stream)))
        status = MARA_ERR_HIP;
    else if (!hip_ok(hipStreamSynchronize(b.stream))) status = MARA_ERR_HIP;
    if (status) return status;
    return fill_report(&b, report);
}

// ===========================================================================
// Block groups: one mesh split into blocks on one device.  The single-device
// half of BlockDecomposition (src/BlockDecomposition.cpp): the blocks take
// the partition of every axis and run their stages on streams of their own,
// and after each stage the group's stream fills their guards from the neighbours' edge
// layers in BoundaryCondition::applySimple order (axis 0, 1, 2; left, then
// right), with the physical condition at the mesh's walls.
// ===========================================================================
struct mara_group
{
    mara_desc d;                      // the whole mesh
    int dims[3] = {1, 1, 1};
    int count = 0;
    hipStream_t stream = nullptr;
    std::vector<mara_block*> blocks;  // row-major over the block coordinates, axis 2 fastest
    std::vector<std::array<int, 3>> start, cells;
    hipEvent_t fork = nullptr;        // recorded on the group's stream, awaited by every block's stream
    std::vector<hipEvent_t> join;     // recorded on each block's stream, awaited by the group's stream
    HaloEntry* halo = nullptr;        // device tables [buffer role][axis][block]
    long long halo_threads[3] = {0, 0, 0};  // largest entry count per axis
    int roles = 2;                    // Pa, Pb (and Pc on the three-buffer RK3 schedule)
    Ctl** ctls = nullptr;             // device array of the members' control blocks
    Ctl* ctl_copy = nullptr;          // device copies of them, gathered for one transfer
    Ctl* ctl_copy_host = nullptr;     // pinned host copy of ctl_copy
};

namespace {

// BlockDecomposition::partition and ::startIndex (src/BlockDecomposition.cpp:187-202):
// the first n % p partitions hold one element more than the others.
void partition_of(int n, int p, int i, int* count, int* start)
{
    const int q = n / p, r = n % p;
    *count = q + (i < r ? 1 : 0);
    const int big = i < r ? i : r;  // partitions of size q + 1 before i
    *start = (q + 1) * big + q * (i - big);
}

int group_coord(const mara_group* g, int m, int axis)
{
    const int c2 = m % g->dims[2];
    const int c1 = (m / g->dims[2]) % g->dims[1];
    const int c0 = m / (g->dims[2] * g->dims[1]);
    return axis == 0 ? c0 : axis == 1 ? c1 : c2;
}

int group_index(const mara_group* g, int c0, int c1, int c2) { return (c0 * g->dims[1] + c1) * g->dims[2] + c2; }

// the buffer of block b in role r: 0 Pa, 1 Pb, 2 Pc
double* role_buffer(mara_block* b, int r) { return r == 0 ? b->Pa : r == 1 ? b->Pb : b->Pc; }

int role_of(const mara_block* b, const double* P) { return P == b->Pa ? 0 : P == b->Pb ? 1 : 2; }

// The blocks' streams wait for the work so far on the group's stream ...
int group_fork(mara_group* g)
{
    HIP_TRY(hipEventRecord(g->fork, g->stream));
    for (mara_block* b : g->blocks) HIP_TRY(hipStreamWaitEvent(b->stream, g->fork, 0));
    return MARA_OK;
}

// ... and the group's stream for the work so far on every block's stream.
int group_join(mara_group* g)
{
    for (int m = 0; m < g->count; ++m)
    {
        HIP_TRY(hipEventRecord(g->join[m], g->blocks[m]->stream));
        HIP_TRY(hipStreamWaitEvent(g->stream, g->join[m], 0));
    }
    return MARA_OK;
}

void group_exchange(mara_group* g, int role)
{
    const int nq = g->blocks[0]->G.nq;
    for (int axis = 0; axis < 3; ++axis)
    {
        if (g->d.guards[axis] == 0 || g->halo_threads[axis] == 0) continue;
        const dim3 grid((unsigned)((g->halo_threads[axis] + kThreads - 1) / kThreads), (unsigned)g->count);
        k_halo<<<grid, kThreads, 0, g->stream>>>(g->halo + ((long long)role * 3 + axis) * g->count, axis, nq);
    }
}

// after group_join: the members' control blocks to the host
int read_group_ctls(mara_group* g)
{
    constexpr int W = sizeof(Ctl) / sizeof(unsigned long long);
    k_ctl_gather<<<(g->count * W + 255) / 256, 256, 0, g->stream>>>(g->ctls, g->count, g->ctl_copy);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(g->ctl_copy_host, g->ctl_copy, g->count * sizeof(Ctl), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    for (int m = 0; m < g->count; ++m) std::memcpy(g->blocks[m]->ctl_host, &g->ctl_copy_host[m], sizeof(Ctl));
    return MARA_OK;
}

// the members' reports as one: counts summed, the first failure the one with
// the smallest linear index of the whole mesh (i slowest)
int group_report(mara_group* g, mara_report* r)
{
    int status = MARA_OK;
    mara_report all = {};
    all.first_failure[0] = all.first_failure[1] = all.first_failure[2] = -1;
    unsigned long long best = ~0ull;
    for (int m = 0; m < g->count; ++m)
    {
        mara_report one;
        const int s = fill_report(g->blocks[m], &one);
        // ctl_status's precedence over the members
        if (s == MARA_ERR_SUPERLUMINAL || (s == MARA_ERR_BAD_EIGENVALUES && status != MARA_ERR_SUPERLUMINAL)
            || (s == MARA_ERR_INVERSION_FAILURE && status == MARA_OK))
            status = s;
        all.c2p_failures += one.c2p_failures;
        all.floor_applied += one.floor_applied;
        all.eigen_failures += one.eigen_failures;
        all.superluminal += one.superluminal;
        if (one.c2p_failures)
        {
            int f[3];
            for (int a = 0; a < 3; ++a) f[a] = one.first_failure[a] + g->start[m][a];
            const unsigned long long lin =
                ((unsigned long long)f[0] * g->d.cells[1] + f[1]) * g->d.cells[2] + f[2];
            if (lin < best)
            {
                best = lin;
                for (int a = 0; a < 3; ++a) all.first_failure[a] = f[a];
            }
        }
    }
    all.status = status;
    if (r) *r = all;
    return status;
}

// Padded local extent along an axis.
inline int padded(const mara_group* g, int m, int a) { return g->cells[m][a] + 2 * g->d.guards[a]; }

} // namespace

extern "C" {

int mara_partition(int n, int p, int which, int* count, int* start)
{
    if (!count || !start || n < 1 || p < 1 || which < 0 || which >= p) return MARA_ERR_INVALID_ARGUMENT;
    partition_of(n, p, which, count, start);
    return MARA_OK;
}

int mara_group_destroy(mara_group* g)
{
    if (!g) return MARA_ERR_INVALID_ARGUMENT;
    if (g->stream) hipStreamSynchronize(g->stream);
    for (mara_block* b : g->blocks)
        if (b) mara_block_destroy(b);
    if (g->fork) hipEventDestroy(g->fork);
    for (hipEvent_t e : g->join)
        if (e) hipEventDestroy(e);
    hipFree(g->halo);
    hipFree(g->ctls);
    hipFree(g->ctl_copy);
    if (g->ctl_copy_host) hipHostFree(g->ctl_copy_host);
    if (g->stream) hipStreamDestroy(g->stream);
    delete g;
    return MARA_OK;
}

int mara_group_create(const mara_desc* d, const int* dims, mara_group** out)
{
    if (!out || !dims) return MARA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int v = validate(d);
    if (v) return v;
    long long total = 1;
    for (int a = 0; a < 3; ++a)
    {
        if (dims[a] < 1) return MARA_ERR_INVALID_ARGUMENT;
        if (dims[a] > 1 && d->cells[a] == 1) return MARA_ERR_INVALID_ARGUMENT;  // a flat axis is not split
        total *= dims[a];
    }
    if (total > MARA_GROUP_MAX_BLOCKS) return MARA_ERR_CAPACITY;
    // flux-CT needs the face fluxes of the neighbours' edges inside a stage,
    // which the guard exchange between stages does not carry; the blocks
    // of a spherical mesh would need their own edge tables
    if (d->law != MARA_LAW_HYDRO && d->field_ct) return MARA_ERR_UNSUPPORTED;
    if (d->geometry != MARA_GEOM_CARTESIAN) return MARA_ERR_UNSUPPORTED;
    for (int a = 0; a < 3; ++a)
    {
        // every block keeps at least two cells and the guard depth along a
        // split axis, so all blocks take the same stage path and every
        // exchange reads interior layers
        const int smallest = d->cells[a] / dims[a];
        if (dims[a] > 1 && (smallest < 2 || smallest < d->guards[a])) return MARA_ERR_UNSUPPORTED;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MARA_ERR_NO_DEVICE;

    mara_group* g = new (std::nothrow) mara_group();
    if (!g) return MARA_ERR_HIP;
    g->d = *d;
    for (int a = 0; a < 3; ++a) g->dims[a] = dims[a];
    g->count = (int)total;
    auto fail = [&](int code) { mara_group_destroy(g); return code; };
    if (!hip_ok(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipEventCreateWithFlags(&g->fork, hipEventDisableTiming))) return fail(MARA_ERR_HIP);
    g->join.assign(g->count, nullptr);
    for (hipEvent_t& e : g->join)
        if (!hip_ok(hipEventCreateWithFlags(&e, hipEventDisableTiming))) return fail(MARA_ERR_HIP);

    // the whole mesh's cell sizes, which every block keeps exactly
    const Geom whole = make_geom(*d);
    g->blocks.assign(g->count, nullptr);
    g->start.resize(g->count);
    g->cells.resize(g->count);
    for (int m = 0; m < g->count; ++m)
    {
        mara_desc local = *d;
        for (int a = 0; a < 3; ++a)
        {
            int n, s;
            partition_of(d->cells[a], dims[a], group_coord(g, m, a), &n, &s);
            g->start[m][a] = s;
            g->cells[m][a] = n;
            local.cells[a] = n;
            // CartesianMeshGeometry of the block (BlockDecomposition::decompose,
            // src/BlockDecomposition.cpp:141-159): its own lower and upper
            const double dx = (d->upper[a] - d->lower[a]) / d->cells[a];
            local.lower[a] = d->lower[a] + s * dx;
            local.upper[a] = d->lower[a] + (s + n) * dx;
        }
        // each block on a stream of its own, so the blocks' stages overlap
        int e = mara_block_create(&local, &g->blocks[m]);
        if (e) return fail(e);
        mara_block* b = g->blocks[m];
        b->in_group = true;
        b->graphs = false;
        for (int a = 0; a < 3; ++a) b->G.dA[a] = whole.dA[a];
        b->G.dV = whole.dV;
        b->G.Lmin = whole.Lmin;
    }
    const mara_block* b0 = g->blocks[0];
    g->roles = b0->Pc ? 3 : 2;

    // exchange tables, one per buffer role and axis
    std::vector<HaloEntry> tab((size_t)g->roles * 3 * g->count);
    for (int r = 0; r < g->roles; ++r)
        for (int axis = 0; axis < 3; ++axis)
            for (int m = 0; m < g->count; ++m)
            {
                HaloEntry& e = tab[((size_t)r * 3 + axis) * g->count + m];
                std::memset(&e, 0, sizeof e);
                mara_block* b = g->blocks[m];
                e.P = role_buffer(b, r);
                e.fs = b->G.fs;
                for (int a = 0; a < 3; ++a) e.N[a] = b->G.N[a];
                e.g = d->guards[axis];
                const int t1 = axis == 0 ? 1 : 0, t2 = axis == 2 ? 1 : 2;
                e.count = 2ll * e.g * e.N[t1] * e.N[t2];
                g->halo_threads[axis] = std::max(g->halo_threads[axis], e.count);
                int c[3] = {group_coord(g, m, 0), group_coord(g, m, 1), group_coord(g, m, 2)};
                for (int side = 0; side < 2; ++side)
                {
                    const bool wall = side == 0 ? c[axis] == 0 : c[axis] == dims[axis] - 1;
                    if (wall && d->bc != MARA_BC_PERIODIC)
                    {
                        e.kind[side] = d->bc;  // outflow 1, reflecting 2
                        continue;
                    }
                    int nc[3] = {c[0], c[1], c[2]};
                    nc[axis] = (c[axis] + (side == 0 ? dims[axis] - 1 : 1)) % dims[axis];
                    mara_block* nb = g->blocks[group_index(g, nc[0], nc[1], nc[2])];
                    e.kind[side] = 0;
                    e.src[side] = role_buffer(nb, r);
                    e.src_fs[side] = nb->G.fs;
                    e.src_N[side] = nb->G.N[axis];
                    if (nb->G.N[t1] != b->G.N[t1] || nb->G.N[t2] != b->G.N[t2]) return fail(MARA_ERR_INVALID_ARGUMENT);
                }
            }
    const size_t tb = tab.size() * sizeof(HaloEntry);
    if (!hip_ok(hipMalloc(&g->halo, tb))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipMemcpy(g->halo, tab.data(), tb, hipMemcpyHostToDevice))) return fail(MARA_ERR_HIP);
    // the members' control blocks, addressed together from the group's stream
    std::vector<Ctl*> ctls(g->count);
    for (int m = 0; m < g->count; ++m) ctls[m] = g->blocks[m]->ctl;
    if (!hip_ok(hipMalloc(&g->ctls, g->count * sizeof(Ctl*)))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipMemcpy(g->ctls, ctls.data(), g->count * sizeof(Ctl*), hipMemcpyHostToDevice)))
        return fail(MARA_ERR_HIP);
    if (!hip_ok(hipMalloc(&g->ctl_copy, g->count * sizeof(Ctl)))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipHostMalloc((void**)&g->ctl_copy_host, g->count * sizeof(Ctl), hipHostMallocDefault)))
        return fail(MARA_ERR_HIP);
    *out = g;
    return MARA_OK;
}

int mara_group_num_blocks(const mara_group* g) { return g ? g->count : -1; }

int mara_group_block_region(const mara_group* g, int index, int* start, int* cells)
{
    if (!g || !start || !cells || index < 0 || index >= g->count) return MARA_ERR_INVALID_ARGUMENT;
    for (int a = 0; a < 3; ++a)
    {
        start[a] = g->start[index][a];
        cells[a] = g->cells[index][a];
    }
    return MARA_OK;
}

long long mara_group_device_bytes(const mara_group* g)
{
    if (!g) return -1;
    long long s = (long long)g->roles * 3 * g->count * sizeof(HaloEntry);
    s += (long long)g->count * (sizeof(Ctl*) + sizeof(Ctl));  // the control-block table and copies
    for (const mara_block* b : g->blocks) s += b->bytes;
    return s;
}

int mara_group_upload_prim(mara_group* g, const double* P)
{
    if (!g || !P) return MARA_ERR_INVALID_ARGUMENT;
    const int nq = g->blocks[0]->G.nq;
    const int GN1 = g->d.cells[1] + 2 * g->d.guards[1], GN2 = g->d.cells[2] + 2 * g->d.guards[2];
    std::vector<double> local;
    for (int m = 0; m < g->count; ++m)
    {
        // the block's window of the whole array, its guards included: padded
        // index l of the block is padded index start + l of the mesh
        const int N0 = padded(g, m, 0), N1 = padded(g, m, 1), N2 = padded(g, m, 2);
        local.resize((size_t)N0 * N1 * N2 * nq);
        const auto& s = g->start[m];
        for (int i = 0; i < N0; ++i)
            for (int j = 0; j < N1; ++j)
            {
                const double* src = P + (((size_t)(s[0] + i) * GN1 + (s[1] + j)) * GN2 + s[2]) * nq;
                std::memcpy(&local[(((size_t)i * N1 + j) * N2) * nq], src, (size_t)N2 * nq * sizeof(double));
            }
        int e = mara_block_upload_prim(g->blocks[m], local.data());
        if (e) return e;
    }
    return MARA_OK;
}

int mara_group_download_prim(mara_group* g, double* P)
{
    if (!g || !P) return MARA_ERR_INVALID_ARGUMENT;
    const int nq = g->blocks[0]->G.nq;
    const int GN1 = g->d.cells[1] + 2 * g->d.guards[1], GN2 = g->d.cells[2] + 2 * g->d.guards[2];
    std::vector<double> local;
    for (int m = 0; m < g->count; ++m)
    {
        const int N0 = padded(g, m, 0), N1 = padded(g, m, 1), N2 = padded(g, m, 2);
        local.resize((size_t)N0 * N1 * N2 * nq);
        int e = mara_block_download_prim(g->blocks[m], local.data());
        if (e) return e;
        // the block's interior, and its guards where they are the mesh's
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a)
        {
            const int c = group_coord(g, m, a);
            lo[a] = c == 0 ? 0 : g->d.guards[a];
            hi[a] = c == g->dims[a] - 1 ? padded(g, m, a) : g->d.guards[a] + g->cells[m][a];
        }
        const auto& s = g->start[m];
        for (int i = lo[0]; i < hi[0]; ++i)
            for (int j = lo[1]; j < hi[1]; ++j)
            {
                double* dst = P + (((size_t)(s[0] + i) * GN1 + (s[1] + j)) * GN2 + s[2] + lo[2]) * nq;
                std::memcpy(dst, &local[(((size_t)i * N1 + j) * N2 + lo[2]) * nq],
                            (size_t)(hi[2] - lo[2]) * nq * sizeof(double));
            }
    }
    return MARA_OK;
}

int mara_group_apply_bc(mara_group* g)
{
    if (!g) return MARA_ERR_INVALID_ARGUMENT;
    int e = group_join(g);
    if (e) return e;
    group_exchange(g, 0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(g->stream));
    return MARA_OK;
}

int mara_group_advance(mara_group* g, double t0, double dt, mara_report* report)
{
    (void)t0;
    if (!g) return MARA_ERR_INVALID_ARGUMENT;
    // per stage: every block's stage on its own stream, then the exchange on
    // the group's stream once all of them are done
    k_ctl_begin_advance_n<<<(g->count + 63) / 64, 64, 0, g->stream>>>(g->ctls, g->count, dt);
    int e = group_fork(g);
    if (e) return e;
    const int rk_order = g->d.rk_order;
    for (int rk = 0; rk < rk_order; ++rk)
    {
        if (rk > 0 && (e = group_fork(g))) return e;
        for (mara_block* b : g->blocks) launch_stage(b, rk, 0);
        if ((e = group_join(g))) return e;
        double* Pin;
        double* Pout;
        stage_io(g->blocks[0], rk, &Pin, &Pout);
        group_exchange(g, role_of(g->blocks[0], Pout));
    }
    HIP_TRY(hipGetLastError());
    e = read_group_ctls(g);
    if (e) return e;
    return group_report(g, report);
}

int mara_group_courant_dt(mara_group* g, double* dt)
{
    if (!g || !dt) return MARA_ERR_INVALID_ARGUMENT;
    int e = group_fork(g);
    if (e) return e;
    for (mara_block* b : g->blocks)
    {
        k_ctl_begin_courant<<<1, 1, 0, b->stream>>>(b->ctl);
        launch_courant(b);
        k_ctl_fold_bins<<<1, 1, 0, b->stream>>>(b->ctl);
    }
    if ((e = group_join(g))) return e;
    HIP_TRY(hipGetLastError());
    e = read_group_ctls(g);
    if (e) return e;
    double best = 1.79769313486231570815e+308;
    int status = MARA_OK;
    for (mara_block* b : g->blocks)
    {
        union { unsigned long long u; double d; } cv;
        cv.u = b->ctl_host->courant_bits;
        best = std::min(best, cv.d);
        if (b->ctl_host->superluminal) status = MARA_ERR_SUPERLUMINAL;
        else if (b->ctl_host->eigen_fail && status == MARA_OK) status = MARA_ERR_BAD_EIGENVALUES;
    }
    *dt = best;
    return status;
}

} // extern "C"
