// batch.hpp -- kernels that advance many independent solves in one launch per
// kernel type per RK stage (the serving path).
//
// Each solve ("item") keeps the device layout and the per-zone arithmetic of
// a single block (kernels.hpp): the batched kernels only replace the launch
// geometry.  A work-group takes one tile from a table built on the host; a
// tile names its item and a patch of that item's faces or zones: 256 x 1
// for 1-D items, 64 x 4 for 2-D and 3-D items.  Ghost zones are written by
// the zone update itself (write_guard_images).  Every solve has its own
// control block (dt, clock, stop flag, failure counters, CFL bins), so
// solves never exchange data and a solve's results are bitwise the same
// alone or in any mix.
#pragma once

#include "kernels.hpp"
#include "fused.hpp"
#include "fused_mhd.hpp"
#include "fused_fast.hpp"

namespace mara {

struct BItem
{
    Geom G;
    double* Pa;
    double* Pb;
    double* U;
    void* F[3];
    Ctl* ctl;
    int bc;
    int lane;  // 0: tile kernels on the batch's first stream, 1: pencil kernels on the second
    int cfl;   // arithmetic of the CFL time step of the item's stage kernels (fast64::CflArith)
    double* S; // NewtonianMHD cooling terms of the stage (nine-field items), else null
};

struct BTile
{
    int item;
    int x0, y0, z0;  // absolute index of the tile's first thread
    int nx, ny;      // valid extent of the tile
    int mode;        // 0: 256 x 1, 1: 64 x 4
    int pad;
};

__device__ __forceinline__ bool tile_thread(const BTile& T, int& i, int& j, int& k)
{
    const int lin = threadIdx.x;
    const int dx = T.mode ? (lin & 63) : lin;
    const int dy = T.mode ? (lin >> 6) : 0;
    i = T.x0 + dx;
    j = T.y0 + dy;
    k = T.z0;
    return dx < T.nx && dy < T.ny;
}

template <int Law, int Recon, int Riem, typename R>
static __global__ void __launch_bounds__(256)
kb_flux(const BItem* __restrict__ items, const BTile* __restrict__ tiles, int axis, int rk)
{
    const BTile T = tiles[blockIdx.x];
    const BItem& it = items[T.item];
    if (!it.ctl->active) return;
    int i, j, k;
    if (!tile_thread(T, i, j, k)) return;
    const double* Pin = rk == 0 ? it.Pa : it.Pb;
    flux_face<Law, Recon, Riem, R>(Pin, (R*)it.F[axis], it.G, axis, i, j, k, it.ctl);
}

template <int Law, typename R, bool CT>
static __global__ void __launch_bounds__(256)
kb_update(const BItem* __restrict__ items, const BTile* __restrict__ tiles, double b, int first, int last,
          int courant_out)
{
    __shared__ double wmin[4];
    const BTile T = tiles[blockIdx.x];
    const BItem& it = items[T.item];
    if (!it.ctl->active) return;
    int i, j, k;
    double dtc = 1.79769313486231570815e+308;
    if (tile_thread(T, i, j, k))
        dtc = update_zone<Law, R, CT>(it.Pa, it.U, last ? it.Pa : it.Pb, (const R*)it.F[0], (const R*)it.F[1],
                                      (const R*)it.F[2], it.G, b, first, !last, courant_out, it.ctl, i, j, k, it.S);
    if (courant_out) courant_to_bins(dtc, wmin, 4, threadIdx.x, it.ctl);
}

// The cooling terms of a nine-field item from the stage's primitives, before
// its flux sweeps (k_cooling_source per tile of the update's zones).
static __global__ void __launch_bounds__(256)
kb_cooling(const BItem* __restrict__ items, const BTile* __restrict__ tiles, int rk)
{
    const BTile T = tiles[blockIdx.x];
    const BItem& it = items[T.item];
    if (!it.ctl->active || !it.S) return;
    int i, j, k;
    if (!tile_thread(T, i, j, k)) return;
    const double* Pin = rk == 0 ? it.Pa : it.Pb;
    const long long c = cell_index(it.G, i, j, k);
    const double gm1 = it.G.gamma - 1.0;
    const double q = Pin[4 * it.G.fs + c] / gm1 * it.G.cool;
    it.S[4 * it.G.fs + c] = -q;
    it.S[8 * it.G.fs + c] = q;
}

// 1-D items of an RK2 batch: reconstruction, fluxes and update in one pass
// per tile of PENCIL - 1 zones (stage 1: Pa -> Pb, stage 2: Pb -> Pa).
template <int Law, int Recon, int Riem, typename R, bool CT, int PENCIL>
static __global__ void __launch_bounds__(PENCIL)
kb_pencil(const BItem* __restrict__ items, const BTile* __restrict__ tiles, double b, int first, int last,
          int courant_out)
{
    constexpr int NQ = PencilLds<Law, Recon, PENCIL>::NQ, W = PencilLds<Law, Recon, PENCIL>::W;
    __shared__ R sP[NQ * W];
    __shared__ R sF[NQ * PENCIL];
    __shared__ double wmin[PENCIL / 64];
    const BTile T = tiles[blockIdx.x];
    const BItem& it = items[T.item];
    if (!it.ctl->active) return;
    const double* Pin = first ? it.Pa : it.Pb;
    double* Pout = last ? it.Pa : it.Pb;
    const double dtc = pencil_tile<Law, Recon, Riem, R, CT, PENCIL>(Pin, it.Pa, it.U, Pout, it.G, b, first, !last,
                                                                    courant_out, it.ctl, T.x0, sP, sF, nullptr);
    if (courant_out) courant_to_bins(dtc, wmin, PENCIL / 64, threadIdx.x, it.ctl);
}

// 2-D and 3-D hydro + PLM items of an RK2 batch: the single block's fused
// stage kernel, one tile (x0, y0, z0 = tile coordinates) per work-group;
// launch order is permuted so neighbouring tiles share an XCD.
template <int Riem, typename R, bool Z>
static __global__ void __launch_bounds__(FUSED_TX * PlmTile<R>::TY) __attribute__((amdgpu_waves_per_eu(PlmTile<R>::WAVES, 8)))
kb_fused_plm(const BItem* __restrict__ items, const BTile* __restrict__ tiles, double b, int first, int last,
             int courant_out, int kz)
{
    const BTile T = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    const BItem& it = items[T.item];
    if (!it.ctl->active) return;
    const Geom G = it.G;
    const double* Pin = first ? it.Pa : it.Pb;
    double* Pout = last ? it.Pa : it.Pb;
    fused_plm_tile<Riem, R, Z>(Pin, it.Pa, it.U, Pout, G, b, first, !last, courant_out, it.ctl, T.x0, T.y0, T.z0,
                               kz, false);
}

// The mixed form's 2-D and 3-D hydro + PLM items: the single block's tiles
// of the fast forms with binary32 faces (fast64::fast_plm_tile).
template <int Riem, bool Z>
static __global__ void __launch_bounds__(fast64::FastTile<float>::TX * fast64::FastTile<float>::TY)
    __attribute__((amdgpu_waves_per_eu(fast64::FastTile<float>::WAVES, 8)))
kb_fast_plm(const BItem* __restrict__ items, const BTile* __restrict__ tiles, double b, int first, int last,
            int courant_out, int kz)
{
    const BTile T = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    const BItem& it = items[T.item];
    if (!it.ctl->active) return;
    const Geom G = it.G;
    const double* Pin = first ? it.Pa : it.Pb;
    double* Pout = last ? it.Pa : it.Pb;
    fast64::fast_plm_tile<float, Riem, Z>(Pin, it.Pa, Pout, G, b, first, courant_out, it.ctl, T.x0, T.y0, T.z0, kz,
                                          false);
}

// ---------------------------------------------------------------------------
// Work-queue form of the fused hydro tile kernels (mara_batch_set_persistent):
// a fixed grid of work-groups that loop over the class's tiles instead of
// one work-group per tile.  Work-group b serves XCD x = b % NUM_XCD and takes,
// in order, the tiles xcd_tile gives that XCD (the same tiles and order as
// the one-per-tile launch), numbering them from the XCD's own counter
// q[x]; it returns once the counter passes the XCD's share, which every
// work-group reaches because each claim raises the counter by one.  The
// counters start at zero (kb_queue_reset before each launch, on the same
// stream).
// ---------------------------------------------------------------------------
static __global__ void __launch_bounds__(64) kb_queue_reset(unsigned* __restrict__ q)
{
    if (threadIdx.x < NUM_XCD) q[threadIdx.x] = 0u;
}

// The next tile of this work-group's XCD share, or -1 when the share is
// done; every thread of the work-group calls it.
__device__ __forceinline__ int queue_next_tile(unsigned* __restrict__ q, int count, int* claim)
{
    const int x = (int)(blockIdx.x % NUM_XCD);
    const int per = count / NUM_XCD, extra = count % NUM_XCD;
    const int size = per + (x < extra ? 1 : 0);
    const int start = x * per + (x < extra ? x : extra);
    if (threadIdx.x == 0 && threadIdx.y == 0) *claim = (int)atomicAdd(&q[x], 1u);
    __syncthreads();
    const int l = *claim;
    __syncthreads();  // every thread has read the claim before the next one overwrites it
    return l < size ? start + l : -1;
}

template <int Riem, typename R, bool Z>
static __global__ void __launch_bounds__(FUSED_TX * PlmTile<R>::TY) __attribute__((amdgpu_waves_per_eu(PlmTile<R>::WAVES, 8)))
kbq_fused_plm(const BItem* __restrict__ items, const BTile* __restrict__ tiles, int count, unsigned* __restrict__ q,
              double b, int first, int last, int courant_out, int kz)
{
    __shared__ int claim;
    for (int t = queue_next_tile(q, count, &claim); t >= 0; t = queue_next_tile(q, count, &claim))
    {
        const BTile T = tiles[t];
        const BItem& it = items[T.item];
        if (!it.ctl->active) continue;
        const double* Pin = first ? it.Pa : it.Pb;
        double* Pout = last ? it.Pa : it.Pb;
        fused_plm_tile<Riem, R, Z>(Pin, it.Pa, it.U, Pout, it.G, b, first, !last, courant_out, it.ctl, T.x0, T.y0,
                                   T.z0, kz, false);
    }
}

template <int Riem, bool Z>
static __global__ void __launch_bounds__(fast64::FastTile<float>::TX * fast64::FastTile<float>::TY)
    __attribute__((amdgpu_waves_per_eu(fast64::FastTile<float>::WAVES, 8)))
kbq_fast_plm(const BItem* __restrict__ items, const BTile* __restrict__ tiles, int count, unsigned* __restrict__ q,
             double b, int first, int last, int courant_out, int kz)
{
    __shared__ int claim;
    for (int t = queue_next_tile(q, count, &claim); t >= 0; t = queue_next_tile(q, count, &claim))
    {
        const BTile T = tiles[t];
        const BItem& it = items[T.item];
        if (!it.ctl->active) continue;
        const double* Pin = first ? it.Pa : it.Pb;
        double* Pout = last ? it.Pa : it.Pb;
        fast64::fast_plm_tile<float, Riem, Z>(Pin, it.Pa, Pout, it.G, b, first, courant_out, it.ctl, T.x0, T.y0,
                                              T.z0, kz, false);
    }
}

// 2-D MHD items of an RK2 batch: the single block's fused MHD tiles (fluxes,
// flux-CT and update in one pass, ghost images included), tile (x0, y0).
template <int Recon, typename R, bool CT, int TY>
static __global__ void __launch_bounds__(64 * TY)
kb_mhd2d(const BItem* __restrict__ items, const BTile* __restrict__ tiles, double b, int first, int last,
         int courant_out)
{
    const BTile T = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    const BItem& it = items[T.item];
    if (!it.ctl->active) return;
    const Geom G = it.G;
    const double* Pin = first ? it.Pa : it.Pb;
    double* Pout = last ? it.Pa : it.Pb;
    mhd2d_tile<Recon, R, CT, TY>(Pin, it.Pa, it.U, Pout, G, b, first, !last, courant_out, it.ctl, T.x0, T.y0, nullptr);
}

// Ghost zones of the fused items (their kernel leaves them): tile x0 is the
// first guard-zone index of ghost_zone's enumeration, nx the count.
static __global__ void __launch_bounds__(256)
kb_ghost(const BItem* __restrict__ items, const BTile* __restrict__ tiles, int last)
{
    const BTile T = tiles[blockIdx.x];
    const BItem& it = items[T.item];
    if (!it.ctl->active) return;
    if ((int)threadIdx.x >= T.nx) return;
    ghost_zone(last ? it.Pa : it.Pb, it.G, (long long)T.x0 + threadIdx.x);
}

template <int Law>
static __global__ void __launch_bounds__(256)
kb_courant(const BItem* __restrict__ items, const BTile* __restrict__ tiles)
{
    __shared__ double wmin[4];
    const BTile T = tiles[blockIdx.x];
    const BItem& it = items[T.item];
    if (!it.ctl->active) return;
    int i, j, k;
    double dtc = 1.79769313486231570815e+308;
    int err = 0;
    if (tile_thread(T, i, j, k))
        dtc = fast64::zone_courant_form<Law>(it.cfl, it.Pa, it.G, cell_index(it.G, i, j, k), it.G.Lmin, err);
    if (err & ERR_SUPERLUMINAL) atomicAdd(&it.ctl->superluminal, 1ull);
    if (err & ERR_EIGEN) atomicAdd(&it.ctl->eigen_fail, 1ull);
    courant_to_bins(dtc, wmin, 4, threadIdx.x, it.ctl);
}

// Per-item step start, one wavefront per item (4 items per work-group of
// 256): the stop condition of k_ctl_step_start, then the CFL bins folded by
// the wavefront and dt = min(cfl * courant, t_final - t).
static __global__ void __launch_bounds__(256) kb_step_start(const BItem* __restrict__ items, int n, int lane)
{
    const int w = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int l = (int)(threadIdx.x & 63);
    if (w >= n || items[w].lane != lane) return;
    Ctl* ctl = items[w].ctl;
    step_start_wave(ctl, l, ctl->active, ctl->t, ctl->steps);
}

// kb_step_end of one step and kb_step_start of the next for every solve of
// a lane, one wavefront per solve: every lane of the wave forms the advanced
// clock from the same loads, lane 0 stores it, the start reads it from
// registers (k_ctl_step_next per solve).
static __global__ void __launch_bounds__(256) kb_step_next(const BItem* __restrict__ items, int n, int lane)
{
    const int w = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int l = (int)(threadIdx.x & 63);
    if (w >= n || items[w].lane != lane) return;
    Ctl* ctl = items[w].ctl;
    const int active = ctl->active;
    double t = ctl->t;
    int steps = ctl->steps;
    if (active)
    {
        t += ctl->dt;
        steps += 1;
        if (l == 0)
        {
            ctl->t = t;
            ctl->steps = steps;
        }
    }
    step_start_wave(ctl, l, active, t, steps);
}

static __global__ void kb_step_end(const BItem* __restrict__ items, int n, int lane)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n || items[t].lane != lane) return;
    Ctl* ctl = items[t].ctl;
    if (!ctl->active) return;
    ctl->t += ctl->dt;
    ctl->steps += 1;
}

} // namespace mara
