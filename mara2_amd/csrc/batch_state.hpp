// batch_state.hpp -- host state of a batch of solves and the launches of one
// RK stage over it, templated on the flux arithmetic R.  The stage launches
// are instantiated for binary64 in batch.hip and for binary32 in mixed.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/mara_gpu.h"
#include "../../include/mara_source.h"
#include "batch.hpp"
#include "host_common.hpp"

namespace mara {
namespace batch {

// first-fit allocator over one device allocation
struct Pool
{
    char* base = nullptr;
    long long size = 0;
    std::map<long long, long long> free_;  // offset -> size

    void init(char* b, long long n) { base = b; size = n; free_.clear(); free_[0] = n; }
    long long available() const
    {
        long long s = 0;
        for (auto& f : free_) s += f.second;
        return s;
    }
    long long alloc(long long n)
    {
        n = host::round_up(n, 256);
        for (auto it = free_.begin(); it != free_.end(); ++it)
        {
            if (it->second >= n)
            {
                const long long off = it->first, rest = it->second - n;
                free_.erase(it);
                if (rest > 0) free_[off + n] = rest;
                return off;
            }
        }
        return -1;
    }
    void release(long long off, long long n)
    {
        n = host::round_up(n, 256);
        auto it = free_.emplace(off, n).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first)
        {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin())
        {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first)
            {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Slot
{
    bool used = false;
    bool fresh = false;
    bool express = false;  // a long chain moved to the express lane
    mara_desc d;
    int cls = -1;
    long long off = 0, bytes = 0;
    long long ncell = 0;
};

// solve class: everything that selects a kernel template
struct ClassKey
{
    int law, recon, riemann, ct;
    int pencil;  // 1-D items of an RK2 batch: fused pencil tiles
    int fused;   // RK2 batch fused stage tiles: hydro + PLM on 2-D (1) or 3-D (2) meshes, MHD on 2-D (3)
    int lane;    // 0 tile lane, 1 pencil lane, 2 express lane (long chains)
    bool operator<(const ClassKey& o) const
    {
        return std::tie(law, recon, riemann, ct, pencil, fused, lane)
               < std::tie(o.law, o.recon, o.riemann, o.ct, o.pencil, o.fused, o.lane);
    }
};

struct Segment { int offset, count; };

} // namespace batch
} // namespace mara

struct mara_batch
{
    int rk_order = 2;
    int precision = MARA_PREC_FP64;
    int max_solves = 0;
    // faces per pencil tile of the 1-D solves: 128, the shortest, is the
    // fastest once the batch fills the device (a saturated 1024-request
    // serving run spent 11.4 s in 1-D MHD and hydro pencils of 128 faces,
    // 13.5 s in pencils of 512)
    int pencil = 128;
    int kz = mara::FUSED_KZ;            // z planes per fused tile of the 3-D hydro solves
    int rows = mara::MHD_ROWS_DEFAULT;  // 2-D MHD tiles of 64 x (4 * rows) zones
    int express_max = 0;                // solves the express lane may hold (0: none)
    int persistent = 0;                 // fused hydro tiles from a work queue: work-groups per CU (0: off)
    // source terms of every solve (mara_batch_set_source): the code object's
    // batched kernel and one launch record per slot for each stage input
    // (Pa in a step's first stage, Pb after), zones = 0 in unused slots
    hipModule_t src_module = nullptr;
    hipFunction_t src_fn = nullptr;
    double src_params[MARA_SOURCE_MAX_PARAMS] = {};
    double src_t[8] = {};
    std::vector<mara_source_args> src_args[2];
    mara_source_args* d_src[2] = {nullptr, nullptr};
    long long src_zones = 0;            // largest interior zone count of a solve admitted with the source
    unsigned* d_queue = nullptr;        // the queue's counters, one set of NUM_XCD per tile lane
    hipStream_t stream = nullptr;   // tile kernels, transfers, control
    hipStream_t stream2 = nullptr;  // pencil kernels of 1-D solves
    hipStream_t stream3 = nullptr;  // express lane: the longest chains of 2-D/3-D solves
    hipStream_t cur = nullptr;      // stream the stage launches go to
    hipEvent_t fork = nullptr, join = nullptr, join3 = nullptr;
    mara::batch::Pool pool;
    char* pool_mem = nullptr;
    mara::BItem* d_items = nullptr;
    mara::Ctl* d_ctl = nullptr;
    unsigned long long* d_bins = nullptr;
    mara::BTile* d_tiles = nullptr;
    int tile_capacity = 0;
    std::vector<mara::BItem> items;
    std::vector<mara::batch::Slot> slots;
    mara::Ctl* h_ctl = nullptr;  // pinned mirror
    std::vector<mara::batch::ClassKey> classes;
    bool dirty = true;
    int n_items = 0;  // highest used slot + 1
    // launch table
    std::vector<mara::BTile> tiles;
    std::map<int, mara::batch::Segment> flux[3];  // class -> segment
    std::map<int, mara::batch::Segment> update;
    std::map<int, mara::batch::Segment> ghost;     // fused classes: ghost-zone tiles
    long long table_version = 0;                   // bumped by every table build
    // HIP graph of one round (a tile-lane step and `ratio` pencil-lane steps),
    // replayed n_steps times by mara_batch_step; recaptured when the table,
    // the slot count, the ratio or the set of busy lanes changes
    bool graphs = true;
    hipGraphExec_t round_exec = nullptr;
    long long round_version = -1;
    int round_n = -1, round_ratio = -1, round_ratio3 = -1, round_lanes = -1;
};

namespace mara {
namespace batch {

// ---- template dispatch over a solve class -------------------------------------
template <int Law, int Recon, int Riem, typename R>
void flux_launch(mara_batch* B, const Segment& s, int axis, int rk)
{
    kb_flux<Law, Recon, Riem, R><<<s.count, 256, 0, B->cur>>>(B->d_items, B->d_tiles + s.offset, axis, rk);
}

template <typename R>
void flux_dispatch(mara_batch* B, const ClassKey& k, const Segment& s, int axis, int rk)
{
#define FL(L, Rc, Rm) if (k.law == L && k.recon == Rc && k.riemann == Rm) { flux_launch<L, Rc, Rm, R>(B, s, axis, rk); return; }
    FL(0, 0, 0) FL(0, 0, 1) FL(0, 1, 0) FL(0, 1, 1) FL(0, 2, 0) FL(0, 2, 1)
    FL(1, 0, 0) FL(1, 1, 0) FL(1, 2, 0)
    FL(2, 0, 0) FL(2, 1, 0) FL(2, 2, 0)
    FL(3, 0, 0) FL(3, 1, 0) FL(3, 2, 0)
#undef FL
}

template <int Law, int Recon, int Riem, typename R, int PL>
void pencil_launch_pl(mara_batch* B, const ClassKey& k, const BTile* t, int count, double b, int first, int last,
                      int cout)
{
    if (k.ct) kb_pencil<Law, Recon, Riem, R, true, PL><<<count, PL, 0, B->cur>>>(B->d_items, t, b, first, last, cout);
    else kb_pencil<Law, Recon, Riem, R, false, PL><<<count, PL, 0, B->cur>>>(B->d_items, t, b, first, last, cout);
}

// the batch's pencil length (one of PENCIL_SHAPES) selects the instantiation
template <int Law, int Recon, int Riem, typename R>
void pencil_launch(mara_batch* B, const ClassKey& k, const BTile* t, int count, double b, int first, int last, int cout)
{
    if constexpr (Law == 2)
    {
        // relativistic solves: the default length only (pencil_length)
        pencil_launch_pl<Law, Recon, Riem, R, PENCIL_DEFAULT>(B, k, t, count, b, first, last, cout);
        return;
    }
    switch (B->pencil)
    {
        case 128: pencil_launch_pl<Law, Recon, Riem, R, 128>(B, k, t, count, b, first, last, cout); break;
        case 512: pencil_launch_pl<Law, Recon, Riem, R, 512>(B, k, t, count, b, first, last, cout); break;
        default: pencil_launch_pl<Law, Recon, Riem, R, 256>(B, k, t, count, b, first, last, cout); break;
    }
}

template <typename R>
void pencil_dispatch(mara_batch* B, const ClassKey& k, const Segment& s, double b, int first, int last, int cout)
{
    const BTile* t = B->d_tiles + s.offset;
#define PN(L, Rc, Rm) if (k.law == L && k.recon == Rc && k.riemann == Rm) { pencil_launch<L, Rc, Rm, R>(B, k, t, s.count, b, first, last, cout); return; }
    PN(0, 0, 0) PN(0, 0, 1) PN(0, 1, 0) PN(0, 1, 1) PN(0, 2, 0) PN(0, 2, 1)
    PN(1, 0, 0) PN(1, 1, 0) PN(1, 2, 0)
    PN(2, 0, 0) PN(2, 1, 0) PN(2, 2, 0)
#undef PN
}

template <int Riem, typename R, bool Z>
void fused_launch(mara_batch* B, const BTile* t, int count, double b, int first, int last, int cout)
{
    // the mixed form runs the single block's tiles of the fast forms (the
    // same 64 x 8 tile shape as PlmTile<float>), the binary64 form the plain ones
    if (B->persistent)
    {
        // a fixed grid drawing the same tiles from per-XCD counters
        unsigned* q = B->d_queue + (B->cur == B->stream3 ? NUM_XCD : 0);
        const int wgs = 256 * B->persistent;
        const int grid = std::max(NUM_XCD, std::min(count + NUM_XCD - 1, wgs) / NUM_XCD * NUM_XCD);
        kb_queue_reset<<<1, 64, 0, B->cur>>>(q);
        if constexpr (sizeof(R) == 4)
            kbq_fast_plm<Riem, Z><<<grid, dim3(64, fast64::FastTile<float>::TY), 0, B->cur>>>(
                B->d_items, t, count, q, b, first, last, cout, B->kz);
        else
            kbq_fused_plm<Riem, R, Z><<<grid, dim3(FUSED_TX, PlmTile<R>::TY), 0, B->cur>>>(
                B->d_items, t, count, q, b, first, last, cout, B->kz);
        return;
    }
    if constexpr (sizeof(R) == 4)
    {
        static_assert(fast64::FastTile<float>::TY == PlmTile<float>::TY, "the batch's tile table assumes one height");
        kb_fast_plm<Riem, Z><<<count, dim3(64, fast64::FastTile<float>::TY), 0, B->cur>>>(B->d_items, t, b, first,
                                                                                       last, cout, B->kz);
    }
    else
        kb_fused_plm<Riem, R, Z><<<count, dim3(FUSED_TX, PlmTile<R>::TY), 0, B->cur>>>(B->d_items, t, b, first, last,
                                                                                      cout, B->kz);
}

template <int Recon, typename R>
void mhd2d_launch(mara_batch* B, const ClassKey& k, const BTile* t, int count, double b, int first, int last,
                  int cout)
{
    auto go = [&](auto ty) {
        constexpr int TY = decltype(ty)::value;
        const dim3 block(64, TY);
        if (k.ct) kb_mhd2d<Recon, R, true, TY><<<count, block, 0, B->cur>>>(B->d_items, t, b, first, last, cout);
        else kb_mhd2d<Recon, R, false, TY><<<count, block, 0, B->cur>>>(B->d_items, t, b, first, last, cout);
    };
    switch (B->rows)
    {
        case 1: go(std::integral_constant<int, 4>{}); break;
        case 2: go(std::integral_constant<int, 8>{}); break;
        default: go(std::integral_constant<int, 16>{}); break;
    }
}

template <typename R>
void update_dispatch(mara_batch* B, const ClassKey& k, const Segment& s, double b, int first, int last, int cout)
{
    if (k.fused == 3)
    {
        const BTile* t = B->d_tiles + s.offset;
        if (k.recon == MARA_RECON_PCM) mhd2d_launch<0, R>(B, k, t, s.count, b, first, last, cout);
        else if (k.recon == MARA_RECON_PLM) mhd2d_launch<1, R>(B, k, t, s.count, b, first, last, cout);
        else mhd2d_launch<2, R>(B, k, t, s.count, b, first, last, cout);
        return;
    }
    if (k.pencil)
    {
        pencil_dispatch<R>(B, k, s, b, first, last, cout);
        return;
    }
    if (k.fused)
    {
        const BTile* t = B->d_tiles + s.offset;
        if (k.fused == 2)
        {
            if (k.riemann) fused_launch<1, R, true>(B, t, s.count, b, first, last, cout);
            else fused_launch<0, R, true>(B, t, s.count, b, first, last, cout);
        }
        else
        {
            if (k.riemann) fused_launch<1, R, false>(B, t, s.count, b, first, last, cout);
            else fused_launch<0, R, false>(B, t, s.count, b, first, last, cout);
        }
        return;
    }
    const BTile* t = B->d_tiles + s.offset;
#define UP(L, C) if (k.law == L && k.ct == C) { kb_update<L, R, (C != 0)><<<s.count, 256, 0, B->cur>>>(B->d_items, t, b, first, last, cout); return; }
    UP(0, 0) UP(1, 0) UP(1, 1) UP(2, 0) UP(2, 1) UP(3, 0) UP(3, 1)
#undef UP
}

// The stage's launches for the classes of one lane (0: tile kernels, 1:
// pencil kernels, 2: express tile kernels), on B->cur.
template <typename R>
void run_stage(mara_batch* B, int rk, int lane)
{
    static const double b1[1] = {1.0};
    static const double b2[2] = {1.0, 0.5};
    static const double b3[3] = {1.0, 1. / 4, 2. / 3};
    const double* bb = B->rk_order == 1 ? b1 : B->rk_order == 2 ? b2 : b3;
    const int last = rk == B->rk_order - 1;
    // source terms of the stage's primitives (MeshOperator::generateSourceTerms),
    // every solve in one launch: solve blockIdx.y of the launch records
    if (B->src_fn && lane == 0 && B->n_items > 0)
    {
        mara_source_args* recs = B->d_src[rk == 0 ? 0 : 1];
        void* args[] = {&recs};
        (void)hipModuleLaunchKernel(B->src_fn, (unsigned)((B->src_zones + 255) / 256), (unsigned)B->n_items, 1, 256, 1,
                                    1, 0, B->cur, args, nullptr);
    }
    // NewtonianMHD cooling (ConservationLaw::addSourceTerms) of the stage's
    // primitives, over the update's tiles of the nine-field classes
    for (auto& e : B->update)
        if (B->classes[e.first].lane == lane && B->classes[e.first].law == 3)
            kb_cooling<<<e.second.count, 256, 0, B->cur>>>(B->d_items, B->d_tiles + e.second.offset, rk);
    for (int axis = 0; axis < 3; ++axis)
        for (auto& e : B->flux[axis])
            if (B->classes[e.first].lane == lane) flux_dispatch<R>(B, B->classes[e.first], e.second, axis, rk);
    for (auto& e : B->update)
        if (B->classes[e.first].lane == lane)
            update_dispatch<R>(B, B->classes[e.first], e.second, bb[rk], rk == 0, last, last);
    // the flux-array and pencil updates write ghost zones themselves
    // (write_guard_images); the fused tiles leave them to this pass
    for (auto& e : B->ghost)
        if (B->classes[e.first].lane == lane) kb_ghost<<<e.second.count, 256, 0, B->cur>>>(B->d_items, B->d_tiles + e.second.offset, last);
}


// binary32 stage launches, compiled with contraction in mixed.hip
void run_stage_f32(mara_batch* B, int rk, int lane);

} // namespace batch
} // namespace mara
