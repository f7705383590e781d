// batch.hip -- host side of the batch (serving) entry points of
// include/mara_gpu.h: a pool of independent solves in one device allocation,
// advanced together, one launch per kernel type per RK stage.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <tuple>
#include <vector>

#include "../../include/mara_gpu.h"
#include "batch_state.hpp"

using namespace mara;
using namespace mara::host;
using namespace mara::batch;

#define HIP_TRY(call) do { if (!hip_ok(call)) return MARA_ERR_HIP; } while (0)

namespace {

int class_of(mara_batch* B, const mara_desc& d, bool express = false)
{
    const bool one_d = d.cells[1] == 1 && d.cells[2] == 1;
    const bool rk2 = B->rk_order == 2;
    // the kernel law: 3 for NewtonianMHD with cooling, which takes the
    // flux-array kernels (nine fields and the cooling terms)
    const int law = law_code(d);
    // fused tiles in RK2 batches: hydro + PLM in 2-D (1) and 3-D (2), MHD in 2-D (3)
    // (with a source every solve takes the flux-array kernels, whose update
    // adds S)
    const bool plain = !B->src_fn;
    int fused = 0;
    if (plain && !one_d && rk2 && law == MARA_LAW_HYDRO && d.recon == MARA_RECON_PLM) fused = d.cells[2] > 1 ? 2 : 1;
    else if (plain && !one_d && rk2 && law == MARA_LAW_MHD && d.cells[2] == 1) fused = 3;
    const int pencil = (plain && one_d && rk2 && law != LAW_MHD_COOLING) ? 1 : 0;
    const ClassKey k{law, d.recon, d.riemann, (d.law != MARA_LAW_HYDRO && d.field_ct) ? 1 : 0,
                     pencil, fused, pencil ? 1 : (express ? 2 : 0)};
    for (size_t c = 0; c < B->classes.size(); ++c)
        if (!(B->classes[c] < k) && !(k < B->classes[c])) return (int)c;
    B->classes.push_back(k);
    return (int)B->classes.size() - 1;
}

long long item_bytes(const mara_desc& d, int precision, bool with_source = false)
{
    int N[3];
    for (int a = 0; a < 3; ++a) N[a] = d.cells[a] + 2 * d.guards[a];
    const int nq = law_nq(law_code(d));
    const long long fs = round_up((long long)N[0] * N[1] * N[2], 64);
    const long long ffs = round_up((long long)(N[0] + 1) * (N[1] + 1) * (N[2] + 1), 64);
    const int real = precision == MARA_PREC_MIXED ? 4 : 8;
    long long b = 3 * round_up(nq * fs * 8, 256);
    for (int a = 0; a < 3; ++a) if (N[a] > 1) b += round_up(nq * ffs * real, 256);
    if (nq == 9 || with_source) b += round_up(nq * fs * 8, 256);  // the cooling or source terms
    return b;
}

// tiles covering the box [lo, hi) of one item
void box_tiles(std::vector<BTile>& out, int item, const int lo[3], const int hi[3])
{
    const int mx = hi[0] - lo[0], my = hi[1] - lo[1];
    for (int k = lo[2]; k < hi[2]; ++k)
    {
        if (my == 1)
        {
            for (int x = 0; x < mx; x += 256)
                out.push_back(BTile{item, lo[0] + x, lo[1], k, std::min(256, mx - x), 1, 0, 0});
        }
        else
        {
            for (int y = 0; y < my; y += 4)
                for (int x = 0; x < mx; x += 64)
                    out.push_back(BTile{item, lo[0] + x, lo[1] + y, k, std::min(64, mx - x), std::min(4, my - y), 1, 0});
        }
    }
}

void build_table(mara_batch* B)
{
    B->table_version += 1;
    B->tiles.clear();
    for (int a = 0; a < 3; ++a) B->flux[a].clear();
    B->update.clear();
    B->ghost.clear();
    auto seg_begin = [&]() { return (int)B->tiles.size(); };
    for (size_t c = 0; c < B->classes.size(); ++c)
    {
        if (B->classes[c].fused == 3)
        {
            // 2-D MHD: one 64 x TY tile per work-group; the kernel writes the
            // ghost images itself
            const int TY = mhd_tile_height(B->rows);
            const int s0 = seg_begin();
            for (int n = 0; n < B->n_items; ++n)
            {
                const Slot& sl = B->slots[n];
                if (!sl.used || sl.cls != (int)c) continue;
                const Geom& G = B->items[n].G;
                const int ntx = (G.n[0] + 63) / 64, nty = (G.n[1] + TY - 1) / TY;
                for (int ty = 0; ty < nty; ++ty)
                    for (int tx = 0; tx < ntx; ++tx) B->tiles.push_back(BTile{n, tx, ty, 0, 0, 0, 3, 0});
            }
            if ((int)B->tiles.size() > s0) B->update[(int)c] = Segment{s0, (int)B->tiles.size() - s0};
            continue;
        }
        if (B->classes[c].fused)
        {
            // one tile per FUSED_TX x TY x KZ block of each item, then the
            // ghost-zone tiles of the same items
            const int TY = B->precision == MARA_PREC_MIXED ? PlmTile<float>::TY : PlmTile<double>::TY;
            int s0 = seg_begin();
            for (int n = 0; n < B->n_items; ++n)
            {
                const Slot& sl = B->slots[n];
                if (!sl.used || sl.cls != (int)c) continue;
                const Geom& G = B->items[n].G;
                const int ntx = (G.n[0] + FUSED_TX - 1) / FUSED_TX, nty = (G.n[1] + TY - 1) / TY;
                const int ntz = G.N[2] > 1 ? (G.n[2] + B->kz - 1) / B->kz : 1;
                for (int tz = 0; tz < ntz; ++tz)
                    for (int ty = 0; ty < nty; ++ty)
                        for (int tx = 0; tx < ntx; ++tx) B->tiles.push_back(BTile{n, tx, ty, tz, 0, 0, 3, 0});
            }
            if ((int)B->tiles.size() > s0) B->update[(int)c] = Segment{s0, (int)B->tiles.size() - s0};
            s0 = seg_begin();
            for (int n = 0; n < B->n_items; ++n)
            {
                const Slot& sl = B->slots[n];
                if (!sl.used || sl.cls != (int)c) continue;
                const Geom& G = B->items[n].G;
                const long long total = 2ll * G.g[0] * G.N[1] * G.N[2] + (long long)G.n[0] * 2 * G.g[1] * G.N[2]
                                        + (long long)G.n[0] * G.n[1] * 2 * G.g[2];
                for (long long x = 0; x < total; x += 256)
                    B->tiles.push_back(BTile{n, (int)x, 0, 0, (int)std::min<long long>(256, total - x), 1, 0, 0});
            }
            if ((int)B->tiles.size() > s0) B->ghost[(int)c] = Segment{s0, (int)B->tiles.size() - s0};
            continue;
        }
        if (B->classes[c].pencil)
        {
            // one tile per pl - 1 interior zones of each 1-D item
            const int pl = pencil_length(B->classes[c].law, B->pencil);
            const int s0 = seg_begin();
            for (int n = 0; n < B->n_items; ++n)
            {
                const Slot& sl = B->slots[n];
                if (!sl.used || sl.cls != (int)c) continue;
                const Geom& G = B->items[n].G;
                const int ie = G.N[0] - G.g[0];
                for (int x0 = G.g[0]; x0 < ie; x0 += pl - 1)
                    B->tiles.push_back(BTile{n, x0, 0, 0, std::min(pl - 1, ie - x0), 1, 2, 0});
            }
            if ((int)B->tiles.size() > s0) B->update[(int)c] = Segment{s0, (int)B->tiles.size() - s0};
            continue;
        }
        for (int axis = 0; axis < 3; ++axis)
        {
            const int s0 = seg_begin();
            for (int n = 0; n < B->n_items; ++n)
            {
                const Slot& sl = B->slots[n];
                if (!sl.used || sl.cls != (int)c) continue;
                const Geom& G = B->items[n].G;
                if (G.N[axis] < 2) continue;
                const bool ctring = B->classes[c].ct;
                int lo[3], hi[3];
                for (int e = 0; e < 3; ++e)
                {
                    if (e == axis) { lo[e] = G.g[e]; hi[e] = G.N[e] - G.g[e] + 1; }
                    else
                    {
                        const int ring = (ctring && G.N[e] > 1) ? 1 : 0;
                        lo[e] = G.g[e] - ring;
                        hi[e] = G.N[e] - G.g[e] + ring;
                    }
                }
                box_tiles(B->tiles, n, lo, hi);
            }
            if ((int)B->tiles.size() > s0) B->flux[axis][(int)c] = Segment{s0, (int)B->tiles.size() - s0};
        }
        const int s0 = seg_begin();
        for (int n = 0; n < B->n_items; ++n)
        {
            const Slot& sl = B->slots[n];
            if (!sl.used || sl.cls != (int)c) continue;
            const Geom& G = B->items[n].G;
            int lo[3], hi[3];
            for (int e = 0; e < 3; ++e) { lo[e] = G.g[e]; hi[e] = G.g[e] + G.n[e]; }
            box_tiles(B->tiles, n, lo, hi);
        }
        if ((int)B->tiles.size() > s0) B->update[(int)c] = Segment{s0, (int)B->tiles.size() - s0};
    }
}

__global__ void kb_reset_bins(Ctl* ctl) { reset_bins(ctl); }

int upload_table(mara_batch* B)
{
    if ((int)B->tiles.size() > B->tile_capacity)
    {
        if (B->d_tiles) HIP_TRY(hipFree(B->d_tiles));
        B->tile_capacity = (int)std::max<size_t>(B->tiles.size() * 2, 1024);
        HIP_TRY(hipMalloc(&B->d_tiles, B->tile_capacity * sizeof(BTile)));
    }
    if (!B->tiles.empty())
        HIP_TRY(hipMemcpyAsync(B->d_tiles, B->tiles.data(), B->tiles.size() * sizeof(BTile), hipMemcpyHostToDevice, B->stream));
    if (B->n_items)
        HIP_TRY(hipMemcpyAsync(B->d_items, B->items.data(), B->n_items * sizeof(BItem), hipMemcpyHostToDevice, B->stream));
    return MARA_OK;
}

// a solve stops on the conditions of maraMainLoop (t >= t_final), at its
// step limit, or at the first failure
bool finished(const Ctl& c)
{
    return !c.active || !(c.t < c.t_final) || c.steps >= c.max_steps || ctl_status(c) != MARA_OK;
}

// Long chains.  Every tile-lane solve takes one step per round, so the solve
// with the most steps left sets the batch's makespan while the rest drain.
// A solve whose remaining steps, estimated from its last dt and clock, exceed
// 4x the median of the running tile-lane solves (and 1000) moves to the
// express lane, which takes several of its steps per round on a stream of
// its own; at most express_max at a time (mara_batch_set_express), and a
// solve stays there until it finishes.  A solve's result does not depend on
// its lane (its steps are its own).

void promote_long_chains(mara_batch* B)
{
    if (B->src_fn) return;  // the source launch serves the tile lane's solves
    std::vector<std::pair<double, int>> rem;
    int express = 0;
    for (int k = 0; k < B->n_items; ++k)
    {
        const Slot& sl = B->slots[k];
        const Ctl& c = B->h_ctl[k];
        if (!sl.used || sl.fresh || finished(c) || B->classes[sl.cls].pencil) continue;
        if (sl.express)
        {
            ++express;
            continue;
        }
        const double by_clock = c.dt > 0 ? (c.t_final - c.t) / c.dt : 0.0;
        rem.push_back({std::min(by_clock, (double)(c.max_steps - c.steps)), k});
    }
    if (rem.size() < 2 || express >= B->express_max) return;
    std::vector<std::pair<double, int>> sorted = rem;
    std::nth_element(sorted.begin(), sorted.begin() + sorted.size() / 2, sorted.end());
    const double threshold = std::max(1000.0, 4.0 * sorted[sorted.size() / 2].first);
    std::sort(rem.begin(), rem.end(), [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
        return a.first > b.first;
    });
    for (const auto& e : rem)
    {
        if (express >= B->express_max || e.first <= threshold) break;
        Slot& sl = B->slots[e.second];
        sl.express = true;
        sl.cls = class_of(B, sl.d, true);
        B->items[e.second].lane = B->classes[sl.cls].lane;
        B->dirty = true;
        ++express;
    }
}

int validate_item(mara_batch* B, const mara_desc* d)
{
    const int v = validate(d);
    if (v) return v;
    if (d->rk_order != B->rk_order || d->precision != B->precision) return MARA_ERR_UNSUPPORTED;
    if (d->geometry != MARA_GEOM_CARTESIAN) return MARA_ERR_UNSUPPORTED;
    // the reference adds the cooling to L before the function's S, an order
    // one S array cannot keep (as for a single block)
    if (B->src_fn && law_code(*d) == LAW_MHD_COOLING) return MARA_ERR_UNSUPPORTED;
    return MARA_OK;
}

// The source kernel's launch records of slot n (both stage inputs), from its
// item; zones = 0 leaves the slot out.
void source_records(mara_batch* B, int n, bool used)
{
    const BItem& it = B->items[n];
    for (int r = 0; r < 2; ++r)
    {
        mara_source_args& a = B->src_args[r][n];
        std::memset(&a, 0, sizeof a);
        if (!used) continue;
        a.P = r == 0 ? it.Pa : it.Pb;
        a.S = it.S;
        a.fs = it.G.fs;
        a.nq = it.G.nq;
        const mara_desc& d = B->slots[n].d;
        for (int e = 0; e < 3; ++e)
        {
            a.N[e] = it.G.N[e];
            a.g[e] = it.G.g[e];
            a.n[e] = it.G.n[e];
            a.lower[e] = d.lower[e];
            a.dx[e] = (d.upper[e] - d.lower[e]) / d.cells[e];  // CartesianMeshGeometry
        }
        for (int q = 0; q < 8; ++q) a.t[q] = B->src_t[q];
        for (int q = 0; q < MARA_SOURCE_MAX_PARAMS; ++q) a.params[q] = B->src_params[q];
    }
}

int upload_source_records(mara_batch* B, int first, int count)
{
    for (int r = 0; r < 2; ++r)
        HIP_TRY(hipMemcpyAsync(B->d_src[r] + first, B->src_args[r].data() + first, count * sizeof(mara_source_args),
                               hipMemcpyHostToDevice, B->stream));
    return MARA_OK;
}

} // namespace

extern "C" {

int mara_batch_create(long long hbm_budget, int rk_order, int precision, int max_solves, mara_batch** out)
{
    if (!out || hbm_budget <= 0 || rk_order < 1 || rk_order > 3 || precision < 0 || precision > 1 || max_solves < 1)
        return MARA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MARA_ERR_NO_DEVICE;
    mara_batch* B = new (std::nothrow) mara_batch();
    if (!B) return MARA_ERR_HIP;
    B->rk_order = rk_order;
    B->precision = precision;
    B->rows = MHD_ROWS_DEFAULT;
    B->max_solves = max_solves;
    auto fail = [&](int code) { mara_batch_destroy(B); return code; };
    if (!hip_ok(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipStreamCreateWithFlags(&B->stream2, hipStreamNonBlocking))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipStreamCreateWithFlags(&B->stream3, hipStreamNonBlocking))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipEventCreateWithFlags(&B->join, hipEventDisableTiming))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipEventCreateWithFlags(&B->join3, hipEventDisableTiming))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipMalloc(&B->pool_mem, hbm_budget))) return fail(MARA_ERR_CAPACITY);
    B->pool.init(B->pool_mem, hbm_budget);
    if (!hip_ok(hipMalloc(&B->d_items, max_solves * sizeof(BItem)))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipMalloc(&B->d_ctl, max_solves * sizeof(Ctl)))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipMalloc(&B->d_bins, (size_t)max_solves * CBINS * CBIN_STRIDE * sizeof(unsigned long long)))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipHostMalloc((void**)&B->h_ctl, max_solves * sizeof(Ctl), hipHostMallocDefault))) return fail(MARA_ERR_HIP);
    if (!hip_ok(hipMalloc(&B->d_queue, 2 * NUM_XCD * sizeof(unsigned)))) return fail(MARA_ERR_HIP);
    std::memset(B->h_ctl, 0, max_solves * sizeof(Ctl));
    B->items.resize(max_solves);
    B->slots.resize(max_solves);
    *out = B;
    return MARA_OK;
}

int mara_batch_destroy(mara_batch* B)
{
    if (!B) return MARA_ERR_INVALID_ARGUMENT;
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    if (B->stream2) (void)hipStreamSynchronize(B->stream2);
    if (B->stream3) (void)hipStreamSynchronize(B->stream3);
    (void)hipFree(B->pool_mem);
    (void)hipFree(B->d_items);
    (void)hipFree(B->d_ctl);
    (void)hipFree(B->d_bins);
    (void)hipFree(B->d_tiles);
    (void)hipFree(B->d_queue);
    (void)hipFree(B->d_src[0]);
    (void)hipFree(B->d_src[1]);
    if (B->src_module) (void)hipModuleUnload(B->src_module);
    if (B->round_exec) (void)hipGraphExecDestroy(B->round_exec);
    if (B->h_ctl) (void)hipHostFree(B->h_ctl);
    if (B->stream) (void)hipStreamDestroy(B->stream);
    if (B->stream2) (void)hipStreamDestroy(B->stream2);
    if (B->stream3) (void)hipStreamDestroy(B->stream3);
    if (B->fork) (void)hipEventDestroy(B->fork);
    if (B->join) (void)hipEventDestroy(B->join);
    if (B->join3) (void)hipEventDestroy(B->join3);
    delete B;
    return MARA_OK;
}

int mara_batch_set_source(mara_batch* B, const char* code_object, const char* kernel, const double* params,
                          int n_params)
{
    if (!B || n_params < 0 || n_params > MARA_SOURCE_MAX_PARAMS || (n_params > 0 && !params)
        || (code_object && !kernel))
        return MARA_ERR_INVALID_ARGUMENT;
    // the solves' layout (S arrays, flux-array classes) is fixed at admission
    for (int s = 0; s < B->max_solves; ++s)
        if (B->slots[s].used) return MARA_ERR_INVALID_ARGUMENT;
    HIP_TRY(hipStreamSynchronize(B->stream));
    if (B->src_module) (void)hipModuleUnload(B->src_module);
    B->src_module = nullptr;
    B->src_fn = nullptr;
    B->src_zones = 0;
    B->classes.clear();  // class keys depend on whether a source is set
    B->dirty = true;
    if (B->round_exec)
    {
        (void)hipGraphExecDestroy(B->round_exec);
        B->round_exec = nullptr;
    }
    if (!code_object) return MARA_OK;
    if (hipModuleLoad(&B->src_module, code_object) != hipSuccess)
    {
        std::fprintf(stderr, "mara_gpu: cannot load source code object %s\n", code_object);
        (void)hipGetLastError();
        B->src_module = nullptr;
        return MARA_ERR_INVALID_ARGUMENT;
    }
    // MARA_SOURCE_KERNEL's batched kernel, <kernel>_batch
    char name[256];
    std::snprintf(name, sizeof name, "%s_batch", kernel);
    if (hipModuleGetFunction(&B->src_fn, B->src_module, name) != hipSuccess)
    {
        std::fprintf(stderr, "mara_gpu: no kernel %s in %s\n", name, code_object);
        (void)hipGetLastError();
        (void)hipModuleUnload(B->src_module);
        B->src_module = nullptr;
        B->src_fn = nullptr;
        return MARA_ERR_INVALID_ARGUMENT;
    }
    for (int r = 0; r < 2; ++r)
    {
        if (!B->d_src[r]) HIP_TRY(hipMalloc(&B->d_src[r], B->max_solves * sizeof(mara_source_args)));
        B->src_args[r].assign(B->max_solves, mara_source_args{});
    }
    for (int q = 0; q < MARA_SOURCE_MAX_PARAMS; ++q) B->src_params[q] = q < n_params ? params[q] : 0.0;
    return upload_source_records(B, 0, B->max_solves) ? MARA_ERR_HIP : MARA_OK;
}

int mara_batch_set_source_aux(mara_batch* B, const double* t)
{
    if (!B || !t) return MARA_ERR_INVALID_ARGUMENT;
    for (int q = 0; q < 8; ++q) B->src_t[q] = t[q];
    if (!B->src_fn) return MARA_OK;
    for (int n = 0; n < B->max_solves; ++n)
        for (int r = 0; r < 2; ++r)
            for (int q = 0; q < 8; ++q) B->src_args[r][n].t[q] = t[q];
    int e = upload_source_records(B, 0, B->max_solves);
    if (e) return e;
    HIP_TRY(hipStreamSynchronize(B->stream));
    return MARA_OK;
}

int mara_batch_set_persistent(mara_batch* B, int workgroups_per_cu)
{
    if (!B || workgroups_per_cu < 0 || workgroups_per_cu > 16) return MARA_ERR_INVALID_ARGUMENT;
    if (B->persistent != workgroups_per_cu && B->round_exec)
    {
        // the captured round holds the other launches
        (void)hipGraphExecDestroy(B->round_exec);
        B->round_exec = nullptr;
    }
    B->persistent = workgroups_per_cu;
    return MARA_OK;
}

int mara_batch_set_express(mara_batch* B, int max_solves)
{
    if (!B || max_solves < 0) return MARA_ERR_INVALID_ARGUMENT;
    B->express_max = max_solves;
    return MARA_OK;
}

int mara_batch_lane(const mara_batch* B, int slot, int* lane)
{
    if (!B || !lane || slot < 0 || slot >= B->max_solves || !B->slots[slot].used) return MARA_ERR_INVALID_ARGUMENT;
    *lane = B->items[slot].lane;
    return MARA_OK;
}

int mara_batch_get_shape(const mara_batch* B, mara_shape* shape)
{
    if (!B || !shape) return MARA_ERR_INVALID_ARGUMENT;
    shape->pencil = B->pencil;
    shape->zplanes = B->kz;
    shape->rows = B->rows;
    return MARA_OK;
}

int mara_batch_set_shape(mara_batch* B, const mara_shape* shape)
{
    if (!B || !shape || !valid_shape(*shape)) return MARA_ERR_INVALID_ARGUMENT;
    if (shape->pencil != B->pencil || shape->zplanes != B->kz || shape->rows != B->rows)
        B->dirty = true;  // tiles change
    B->pencil = shape->pencil;
    B->kz = shape->zplanes;
    B->rows = shape->rows;
    return MARA_OK;
}

int mara_batch_set_graphs(mara_batch* B, int enable)
{
    if (!B) return MARA_ERR_INVALID_ARGUMENT;
    B->graphs = enable != 0;
    return MARA_OK;
}

long long mara_batch_free_bytes(const mara_batch* B)
{
    return B ? B->pool.available() : -1;
}

long long mara_batch_item_bytes(const mara_desc* d, int precision)
{
    if (!d) return -1;
    return item_bytes(*d, precision);
}

int mara_batch_admit(mara_batch* B, const mara_desc* d, const double* P, double cfl, double t_final,
                     int max_steps, int* slot)
{
    if (!B || !P || !slot) return MARA_ERR_INVALID_ARGUMENT;
    const int v = validate_item(B, d);
    if (v) return v;
    int n = -1;
    for (int s = 0; s < B->max_solves; ++s)
        if (!B->slots[s].used) { n = s; break; }
    if (n < 0) return MARA_ERR_CAPACITY;
    const long long bytes = item_bytes(*d, B->precision, B->src_fn != nullptr);
    const long long off = B->pool.alloc(bytes);
    if (off < 0) return MARA_ERR_CAPACITY;

    Slot& sl = B->slots[n];
    sl.used = true;
    sl.fresh = true;
    sl.d = *d;
    sl.off = off;
    sl.bytes = bytes;
    sl.cls = class_of(B, *d);
    BItem& it = B->items[n];
    std::memset(&it, 0, sizeof it);
    it.G = make_geom(*d);
    sl.ncell = (long long)it.G.N[0] * it.G.N[1] * it.G.N[2];
    char* p = B->pool.base + off;
    const long long cb = round_up(it.G.nq * it.G.fs * 8, 256);
    it.Pa = (double*)p; p += cb;
    it.Pb = (double*)p; p += cb;
    it.U = (double*)p; p += cb;
    const int real = B->precision == MARA_PREC_MIXED ? 4 : 8;
    for (int a = 0; a < 3; ++a)
        if (it.G.N[a] > 1) { it.F[a] = p; p += round_up(it.G.nq * it.G.ffs * real, 256); }
    it.S = nullptr;
    if (it.G.nq == 9 || B->src_fn)
    {
        // the cooling terms; fields other than the energy and the auxiliary
        // one stay zero
        it.S = (double*)p;
        p += cb;
    }
    it.ctl = B->d_ctl + n;
    it.bc = d->bc;
    it.lane = B->classes[sl.cls].lane;
    // the first step's CFL in the arithmetic of the item's stage kernels:
    // the mixed form's hydro PLM tiles are the fast binary32 ones
    {
        const mara::batch::ClassKey& key = B->classes[sl.cls];
        if (B->precision != MARA_PREC_MIXED) it.cfl = fast64::CFL_EXACT;
        else if (key.fused == 1 || key.fused == 2) it.cfl = fast64::CFL_FAST32;
        else it.cfl = key.law < 2 ? fast64::CFL_MIXED : fast64::CFL_EXACT;
    }
    sl.express = false;

    Ctl& c = B->h_ctl[n];
    std::memset(&c, 0, sizeof c);
    c.t = 0.0;
    c.t_final = t_final;
    c.cfl = cfl;
    c.max_steps = max_steps;
    c.active = 1;
    c.first_fail = ~0ull;
    c.courant_bits = DBL_MAX_BITS;
    c.cbins = B->d_bins + (size_t)n * CBINS * CBIN_STRIDE;
    // a failed transfer gives the slot and the pool range back
    auto undo = [&]() {
        (void)hipStreamSynchronize(B->stream);
        B->pool.release(sl.off, sl.bytes);
        sl.used = false;
        std::memset(&c, 0, sizeof c);
        return MARA_ERR_HIP;
    };
    if (!hip_ok(hipMemcpyAsync(B->d_ctl + n, &c, sizeof(Ctl), hipMemcpyHostToDevice, B->stream))) return undo();
    kb_reset_bins<<<1, 1, 0, B->stream>>>(B->d_ctl + n);
    // primitives: host -> staging (U) -> Pa, guards included; Pb starts as a copy
    if (!hip_ok(hipMemcpyAsync(it.U, P, sl.ncell * it.G.nq * sizeof(double), hipMemcpyHostToDevice, B->stream)))
        return undo();
    k_aos_to_soa<<<(unsigned)((sl.ncell + 255) / 256), 256, 0, B->stream>>>(it.U, it.Pa, it.G);
    if (!hip_ok(hipMemcpyAsync(it.Pb, it.Pa, cb, hipMemcpyDeviceToDevice, B->stream))) return undo();
    if (it.S && !hip_ok(hipMemsetAsync(it.S, 0, cb, B->stream))) return undo();
    if (!hip_ok(hipGetLastError())) return undo();
    if (B->src_fn)
    {
        source_records(B, n, true);
        if (upload_source_records(B, n, 1)) return undo();
        B->src_zones = std::max(B->src_zones, (long long)it.G.n[0] * it.G.n[1] * it.G.n[2]);
        if (B->round_exec)
        {
            // the captured round's source launch covers the old slot count and zone count
            (void)hipGraphExecDestroy(B->round_exec);
            B->round_exec = nullptr;
        }
    }
    // the host buffer may go away after return
    if (!hip_ok(hipStreamSynchronize(B->stream))) return undo();
    if (n + 1 > B->n_items) B->n_items = n + 1;
    B->dirty = true;
    *slot = n;
    return MARA_OK;
}

int mara_batch_step(mara_batch* B, int n_steps, int* n_active)
{
    if (!B || n_steps < 0) return MARA_ERR_INVALID_ARGUMENT;
    promote_long_chains(B);
    if (B->dirty)
    {
        build_table(B);
        int e = upload_table(B);
        if (e) return e;
        B->dirty = false;
    }
    // first CFL of newly admitted solves, grouped by law
    std::vector<BTile> fresh[3];
    for (int n = 0; n < B->n_items; ++n)
    {
        Slot& sl = B->slots[n];
        if (!sl.used || !sl.fresh) continue;
        const Geom& G = B->items[n].G;
        int lo[3], hi[3];
        for (int e = 0; e < 3; ++e) { lo[e] = G.g[e]; hi[e] = G.g[e] + G.n[e]; }
        box_tiles(fresh[sl.d.law], n, lo, hi);  // a cooled MHD solve's CFL is the MHD one
        sl.fresh = false;
    }
    std::vector<BTile> ftiles;
    int fofs[4] = {0, 0, 0, 0};
    for (int l = 0; l < 3; ++l)
    {
        ftiles.insert(ftiles.end(), fresh[l].begin(), fresh[l].end());
        fofs[l + 1] = (int)ftiles.size();
    }
    BTile* d_ftiles = nullptr;
    if (!ftiles.empty())
    {
        HIP_TRY(hipMallocAsync((void**)&d_ftiles, ftiles.size() * sizeof(BTile), B->stream));
        HIP_TRY(hipMemcpyAsync(d_ftiles, ftiles.data(), ftiles.size() * sizeof(BTile), hipMemcpyHostToDevice, B->stream));
        if (fofs[1] > fofs[0]) kb_courant<0><<<fofs[1] - fofs[0], 256, 0, B->stream>>>(B->d_items, d_ftiles + fofs[0]);
        if (fofs[2] > fofs[1]) kb_courant<1><<<fofs[2] - fofs[1], 256, 0, B->stream>>>(B->d_items, d_ftiles + fofs[1]);
        if (fofs[3] > fofs[2]) kb_courant<2><<<fofs[3] - fofs[2], 256, 0, B->stream>>>(B->d_items, d_ftiles + fofs[2]);
    }
    const int n = B->n_items;
    const unsigned ig = (unsigned)((n + 63) / 64);
    // Three lanes: tile kernels (2-D/3-D solves, and 1-D ones outside RK2) on
    // the first stream, pencil kernels of 1-D RK2 solves on the second, so
    // that small 1-D solves are not held to the pace of large ones, and the
    // express lane (promote_long_chains) on the third.  A round is one
    // tile-lane step; the other lanes take `ratio` / `ratio3` steps in it,
    // sized from the tile lane's estimated step time (its active zones at
    // ~5e9 zone-stages per second, plus ~20 us of launches) against theirs.
    long long zones[3] = {0, 0, 0};
    bool busy[3] = {false, false, false};
    for (int k = 0; k < n; ++k)
    {
        if (!B->slots[k].used || finished(B->h_ctl[k])) continue;
        const int l = B->items[k].lane;
        busy[l] = true;
        zones[l] += (long long)B->items[k].G.n[0] * B->items[k].G.n[1] * B->items[k].G.n[2];
    }
    const bool any_tile = busy[0], any_pencil = busy[1], any_express = busy[2];
    auto step_time = [&](int l) { return (double)zones[l] * B->rk_order / 5e9 + 20e-6; };
    const double round_time = any_tile ? step_time(0) : any_express ? step_time(2) : 20e-6;
    int ratio = 1, ratio3 = 1;
    if (any_pencil && (any_tile || any_express)) ratio = (int)std::min(32.0, std::max(1.0, round_time / 20e-6));
    if (any_express && any_tile) ratio3 = (int)std::min(32.0, std::max(1.0, round_time / step_time(2)));
    const unsigned sg = (unsigned)((n + 3) / 4);  // kb_step_start: 4 items per work-group
    // `count` steps of one lane: one kernel ends a step and starts the next
    auto run_lane = [&](int lane, int count) {
        kb_step_start<<<sg, 256, 0, B->cur>>>(B->d_items, n, lane);
        for (int s = 0; s < count; ++s)
        {
            for (int rk = 0; rk < B->rk_order; ++rk)
            {
                if (B->precision == MARA_PREC_MIXED) run_stage_f32(B, rk, lane);
                else run_stage<double>(B, rk, lane);
            }
            if (s + 1 < count) kb_step_next<<<sg, 256, 0, B->cur>>>(B->d_items, n, lane);
            else kb_step_end<<<ig, 64, 0, B->cur>>>(B->d_items, n, lane);
        }
    };
    auto run_round = [&]() {
        if (any_tile || (!any_pencil && !any_express))
        {
            B->cur = B->stream;
            run_lane(0, 1);
        }
        if (any_pencil)
        {
            B->cur = B->stream2;
            run_lane(1, ratio);
        }
        if (any_express)
        {
            B->cur = B->stream3;
            run_lane(2, ratio3);
        }
    };
    const int lanes = (any_tile ? 1 : 0) | (any_pencil ? 2 : 0) | (any_express ? 4 : 0);
    if (B->graphs && n > 0 && n_steps > 1)
    {
        if (!B->round_exec || B->round_version != B->table_version || B->round_n != n || B->round_ratio != ratio
            || B->round_ratio3 != ratio3 || B->round_lanes != lanes)
        {
            if (B->round_exec) HIP_TRY(hipGraphExecDestroy(B->round_exec));
            B->round_exec = nullptr;
            // the round forks the other streams off the first and joins them
            // back, so one replay ends when every lane has
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
            (void)hipEventRecord(B->fork, B->stream);
            (void)hipStreamWaitEvent(B->stream2, B->fork, 0);
            (void)hipStreamWaitEvent(B->stream3, B->fork, 0);
            run_round();
            (void)hipEventRecord(B->join, B->stream2);
            (void)hipEventRecord(B->join3, B->stream3);
            (void)hipStreamWaitEvent(B->stream, B->join, 0);
            (void)hipStreamWaitEvent(B->stream, B->join3, 0);
            const hipError_t launched = hipGetLastError();
            const hipError_t ended = hipStreamEndCapture(B->stream, &g);
            if (!hip_ok(launched) || !hip_ok(ended))
            {
                if (g) (void)hipGraphDestroy(g);
                return MARA_ERR_HIP;
            }
            const hipError_t made = hipGraphInstantiate(&B->round_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (!hip_ok(made))
            {
                B->round_exec = nullptr;
                return MARA_ERR_HIP;
            }
            B->round_version = B->table_version;
            B->round_n = n;
            B->round_ratio = ratio;
            B->round_ratio3 = ratio3;
            B->round_lanes = lanes;
        }
        for (int s = 0; s < n_steps; ++s) HIP_TRY(hipGraphLaunch(B->round_exec, B->stream));
    }
    else
    {
        HIP_TRY(hipEventRecord(B->fork, B->stream));
        HIP_TRY(hipStreamWaitEvent(B->stream2, B->fork, 0));
        HIP_TRY(hipStreamWaitEvent(B->stream3, B->fork, 0));
        for (int s = 0; s < n_steps && n > 0; ++s) run_round();
        HIP_TRY(hipEventRecord(B->join, B->stream2));
        HIP_TRY(hipEventRecord(B->join3, B->stream3));
        HIP_TRY(hipStreamWaitEvent(B->stream, B->join, 0));
        HIP_TRY(hipStreamWaitEvent(B->stream, B->join3, 0));
    }
    HIP_TRY(hipGetLastError());
    if (n) HIP_TRY(hipMemcpyAsync(B->h_ctl, B->d_ctl, n * sizeof(Ctl), hipMemcpyDeviceToHost, B->stream));
    if (d_ftiles) HIP_TRY(hipFreeAsync(d_ftiles, B->stream));
    HIP_TRY(hipStreamSynchronize(B->stream));
    int active = 0;
    for (int k = 0; k < n; ++k)
    {
        const Ctl& c = B->h_ctl[k];
        if (!B->slots[k].used) continue;
        if (!finished(c)) ++active;
    }
    if (n_active) *n_active = active;
    return MARA_OK;
}

int mara_batch_query(mara_batch* B, int slot, int* done, double* t, int* steps, mara_report* r)
{
    if (!B || slot < 0 || slot >= B->max_solves || !B->slots[slot].used) return MARA_ERR_INVALID_ARGUMENT;
    const Ctl& c = B->h_ctl[slot];
    if (done) *done = finished(c) ? 1 : 0;
    if (t) *t = c.t;
    if (steps) *steps = c.steps;
    return fill_report(c, B->items[slot].G, r);
}

int mara_batch_finished(const mara_batch* B, int* slots, int capacity, int* count)
{
    if (!B || !count || capacity < 0 || (capacity > 0 && !slots)) return MARA_ERR_INVALID_ARGUMENT;
    int m = 0;
    for (int n = 0; n < B->n_items && m < capacity; ++n)
        if (B->slots[n].used && finished(B->h_ctl[n])) slots[m++] = n;
    *count = m;
    return MARA_OK;
}

int mara_batch_retire(mara_batch* B, int slot, double* P)
{
    if (!B || slot < 0 || slot >= B->max_solves || !B->slots[slot].used) return MARA_ERR_INVALID_ARGUMENT;
    Slot& sl = B->slots[slot];
    BItem& it = B->items[slot];
    if (P)
    {
        k_soa_to_aos<<<(unsigned)((sl.ncell + 255) / 256), 256, 0, B->stream>>>(it.Pa, it.U, it.G);
        HIP_TRY(hipMemcpyAsync(P, it.U, sl.ncell * it.G.nq * sizeof(double), hipMemcpyDeviceToHost, B->stream));
    }
    // the slot's control block stays inactive until the slot is reused: the
    // per-solve control kernels still visit it while lower slots are in use
    Ctl& c = B->h_ctl[slot];
    std::memset(&c, 0, sizeof c);
    c.cbins = B->d_bins + (size_t)slot * CBINS * CBIN_STRIDE;
    HIP_TRY(hipMemcpyAsync(B->d_ctl + slot, &c, sizeof(Ctl), hipMemcpyHostToDevice, B->stream));
    if (B->src_fn)
    {
        source_records(B, slot, false);
        int e = upload_source_records(B, slot, 1);
        if (e) return e;
    }
    HIP_TRY(hipStreamSynchronize(B->stream));
    B->pool.release(sl.off, sl.bytes);
    sl.used = false;
    while (B->n_items > 0 && !B->slots[B->n_items - 1].used) --B->n_items;
    B->dirty = true;
    return MARA_OK;
}

} // extern "C"
