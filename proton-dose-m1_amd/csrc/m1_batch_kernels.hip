// Batched-beamlet kernels for gfx950: one launch advances the evaluated tiles
// of every beamlet box of the batch.
//
// The per-tile arithmetic is the FP32-bulk stage body of m1_f32_common.h, so
// a beamlet marched in the batch follows the same arithmetic as the same box
// marched alone.  What is batch-specific is the work distribution: a stage
// launch runs one 256-thread workgroup per entry of a compacted, ascending
// list of (box, tile) pairs, and workgroups are mapped so that each of the 8
// XCDs (one L2 each) takes one contiguous slice of the list, keeping the
// halo re-reads of neighbouring tiles in the same L2.
#include <atomic>

#include "m1_f32_common.h"
#include "m1_mlp.h"

// Compiled once per tile shape: M1_TILE_TX / M1_TILE_TY set the shape,
// M1_BATCH_NS the namespace of this copy's launchers and M1_BATCH_TABLE the
// function returning their table (m1_device.h, M1_BATCH_SHAPES).
#ifndef M1_BATCH_NS
#define M1_BATCH_NS batch_16x16x16
#define M1_BATCH_TABLE batch_kernels_16x16x16
#endif

namespace m1 {
namespace M1_BATCH_NS {
namespace {

using namespace f32::detail;
using f32::TileView;

constexpr int XCDS = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin

// List slot of workgroup i: XCD x = i % 8 receives the slots
// [x q + min(x, r), ...) with n = 8 q + r, in order.
__device__ __forceinline__ int xcd_slot(int i, int n)
{
    const int q = n / XCDS, r = n % XCDS;
    const int x = i % XCDS, j = i / XCDS;
    return x * q + min(x, r) + j;
}

struct TileCoord {
    int b, tx, ty, tz;
};

__device__ __forceinline__ TileCoord tile_coord(const BoxDesc& d, int b, int t)
{
    const int local = t - d.tile_off;
    return {b, local / (d.ntz * d.nty), (local / d.ntz) % d.nty, local % d.ntz};
}

// The first relaxation sweep of a box after it is done: the tile's
// quasi-steady cells keep their values (V1, or src on the first sweep, into
// out).
__device__ void relax_copy_tile(const BatchStageArgs& a, int b, int t)
{
    const BoxDesc d = a.box[b];
    const TileCoord tc = tile_coord(d, b, t);
    const long N = long(d.nx) * d.ny * d.nz;
    const int z = tc.tz * TZ + threadIdx.x % TZ;
    const int y = tc.ty * TY + threadIdx.x / TZ;
    if (z >= d.nz || y >= d.ny)
        return;
    const long off4 = 4 * d.cell_off;
    const float* from = (a.V1 ? a.V1 : a.src) + off4;
    float* to = a.out + off4;
    const StepCoef32* c = a.coef + b;
    for (int x = tc.tx * TX; x < min(tc.tx * TX + TX, d.nx); ++x) {
        const long i = (long(x) * d.ny + y) * d.nz + z;
        if (c->dEm[a.ids[d.cell_off + i]] == 0.f)
            continue;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            to[i + q * N] = from[i + q * N];
    }
}

template <int KIND, bool IFAC, bool D>
__device__ __forceinline__ void run_tile(const BatchStageArgs& a, int t)
{
    const int b = a.tile_box[t];
    if constexpr (KIND == STAGE_RELAX) {
        if (a.relax_done_in) {
            // the same decision in every workgroup of the box; a box that has
            // just become done copies its cells into this sweep's output, so
            // that both outputs hold them and later sweeps leave the box alone
            const bool was = a.relax_done_in[b] != 0;
            const bool done = was || __uint_as_float(a.relax_change_prev[b]) < a.relax_tol;
            if (threadIdx.x == 0) {
                a.relax_done_out[b] = done ? 1 : 0;
                a.relax_change_clear[b] = 0u;
            }
            if (done) {
                if (!was)
                    relax_copy_tile(a, b, t);
                return;
            }
        }
    }
    const BoxDesc d = a.box[b];
    const TileCoord tc = tile_coord(d, b, t);
    const long off4 = 4 * d.cell_off;
    const long m2 = long(d.nx) * d.ny;
    float* dose_hi = a.dose_hi;
    float* dose_lo = a.dose_lo;
    if (dose_hi && a.judge && a.judge[b]) {
        dose_hi = a.judge_hi;
        dose_lo = a.judge_lo;
    }
    TileView v{{d.nx, d.ny, d.nz, a.dx, a.dy, a.dz},
               a.src + off4,
               a.U + off4,
               a.V1 ? a.V1 + off4 : nullptr,
               a.out ? a.out + off4 : nullptr,
               dose_hi ? dose_hi + d.cell_off : nullptr,
               dose_lo ? dose_lo + d.cell_off : nullptr,
               nullptr,
               a.ids + d.cell_off,
               a.bnd + 12 * d.col_off + 4 * m2 * a.bnd_stage,
               a.coef + b,
               KIND == STAGE_RELAX
                   ? (a.relax_done_in ? reinterpret_cast<float*>(a.relax_change_cur + b) : nullptr)
                   : (a.tile_max ? a.tile_max + t : nullptr),
               a.tile_sums ? a.tile_sums + 2 * long(t) : nullptr,
               a.tile_face ? a.tile_face + FACE_SLOTS * long(t) : nullptr,
               a.wave_mode,
               tc.tx,
               tc.ty,
               tc.tz,
               a.win ? a.win + off4 : nullptr,
               a.wout ? a.wout + off4 : nullptr};
    stage_tile<KIND, IFAC, D, true>(v);
}

// Launch forms of the stage kernels, one kernel each (a loop around the tile
// body makes stages 1 and 2 slower than one tile per workgroup, so only the
// forms that need a loop have one):
//   LAUNCH_TILE      one workgroup per list entry: the list length is the
//                    grid (a.strided == 0), or (a.strided > 0, an estimate
//                    of it) read from count[0], workgroups past it exiting;
//   LAUNCH_OVERFLOW  the entries past an estimate a.strided, grid-stride
//                    (the grid and a.strided are multiples of 8, so an entry
//                    stays on the XCD of its slot);
//   LAUNCH_QUEUE     a persistent grid pulling entries from the per-XCD
//                    queues: a workgroup of XCD x (workgroups are dispatched
//                    round-robin over the XCDs) drains slice x first,
//                    keeping halo re-reads of neighbouring tiles in that
//                    XCD's L2, then steals from the others.  Queue counters
//                    only grow, so every workgroup ends once all slices are
//                    handed out.
enum StageLaunch : int { LAUNCH_TILE = 0, LAUNCH_OVERFLOW = 1, LAUNCH_QUEUE = 2 };

template <int KIND, bool IFAC, bool D, int L>
__global__ void __launch_bounds__(NT, M1_STAGE_WAVES) stage_kernel(BatchStageArgs a)
{
    if constexpr (L == LAUNCH_TILE) {
        const int n = a.strided ? *a.count : a.n_list;
        const int i = int(blockIdx.x);
        if (i < n)
            run_tile<KIND, IFAC, D>(a, a.list[xcd_slot(i, n)]);
    } else if constexpr (L == LAUNCH_OVERFLOW) {
        const int n = *a.count;
        const int step = int(gridDim.x);
        for (int j = a.strided + int(blockIdx.x); j < n; j += step) {
            run_tile<KIND, IFAC, D>(a, a.list[xcd_slot(j, n)]);
            if (j + step < n)
                __syncthreads();  // the next tile reuses LDS
        }
    } else {
        __shared__ int s_slot;
        const int n = *a.count;
        const int q = n / XCD_QUEUES, r = n % XCD_QUEUES;
        const int home = int(blockIdx.x) % XCD_QUEUES;
        for (;;) {
            if (threadIdx.x == 0) {
                int got = -1;
                for (int k = 0; k < XCD_QUEUES && got < 0; ++k) {
                    const int x = (home + k) % XCD_QUEUES;
                    const int len = q + (x < r ? 1 : 0);
                    const int s = atomicAdd(a.queue + x, 1);
                    if (s < len)
                        got = x * q + min(x, r) + s;
                }
                s_slot = got;
            }
            __syncthreads();
            const int slot = s_slot;
            if (slot < 0)
                break;
            run_tile<KIND, IFAC, D>(a, a.list[slot]);
            __syncthreads();  // the next tile reuses LDS and s_slot
        }
    }
}

// Vacuum-aware CFL of one live tile (energy.py:459-537, 610-636): cell
// activity over the tile and a 3-cell halo along each axis (cells of tiles
// outside the live set hold the zero state and are inactive), then each
// cell's axis speed counts where an active cell lies within 3 cells along
// that axis.  Tiles outside the live set contribute nothing: an active cell
// within 3 cells of one would have kept it in the set, whose reach is 9.
constexpr int CFL_HALO = 3;
constexpr int AX = TX + 2 * CFL_HALO, AY = TY + 2 * CFL_HALO, AZ = TZ + 2 * CFL_HALO;

__device__ void cfl_tile_vacuum_aware(const BatchCflArgs& a, const BoxDesc& d,
                                      const TileCoord& tc, int b, double* red)
{
    __shared__ uint8_t act[AX][AY][AZ];
    const StepCoef& c = a.coef[b];
    const GridDims g{d.nx, d.ny, d.nz, a.dx, a.dy, a.dz};
    const long N = long(d.nx) * d.ny * d.nz;
    const double pin_scale = c.ed[0] * c.inv_alpha, rf = c.rf;
    const uint8_t* ids = a.ids + d.cell_off;
    const double* spatial = a.spatial + d.col_off;
    const float* U = a.U + 4 * d.cell_off;
    const int tid = threadIdx.x;
    const int x0 = tc.tx * TX, y0 = tc.ty * TY, z0 = tc.tz * TZ;
    for (int k = tid; k < AX * AY * AZ; k += NT) {
        const int lz = k % AZ, ly = (k / AZ) % AY, lx = k / (AZ * AY);
        const int out = (lx < CFL_HALO || lx >= TX + CFL_HALO) + (ly < CFL_HALO || ly >= TY + CFL_HALO) +
                        (lz < CFL_HALO || lz >= TZ + CFL_HALO);
        const int x = x0 + lx - CFL_HALO, y = y0 + ly - CFL_HALO, z = z0 + lz - CFL_HALO;
        bool on = false;
        if (out <= 1 && x >= 0 && x < d.nx && y >= 0 && y < d.ny && z >= 0 && z < d.nz) {
            const int ux = tc.tx + (lx < CFL_HALO ? -1 : (lx >= TX + CFL_HALO ? 1 : 0));
            const int uy = tc.ty + (ly < CFL_HALO ? -1 : (ly >= TY + CFL_HALO ? 1 : 0));
            const int uz = tc.tz + (lz < CFL_HALO ? -1 : (lz >= TZ + CFL_HALO ? 1 : 0));
            const int u = d.tile_off + (ux * d.nty + uy) * d.ntz + uz;
            const long i = (long(x) * d.ny + y) * d.nz + z;
            if (a.live[u])
                on = U[i] > a.act_threshold;
            if (z == 0)
                on = on || spatial[long(x) * d.ny + y] * pin_scale > double(a.act_threshold);
        }
        act[lx][ly][lz] = on ? 1 : 0;
    }
    __syncthreads();
    const double eps = c.eig_eps;
    const int tz = tid % TZ, ty = tid / TZ;
    const int z = z0 + tz, y = y0 + ty;
    const int x1 = min(x0 + TX, d.nx);
    double mx = 0.0;
    bool bad = false;
    if (z < d.nz && y < d.ny) {
        for (int x = x0; x < x1; ++x) {
            const int lx = x - x0 + CFL_HALO, ly = ty + CFL_HALO, lz = tz + CFL_HALO;
            bool mxk = false, myk = false, mzk = false;
#pragma unroll
            for (int q = -CFL_HALO; q <= CFL_HALO; ++q) {
                mxk = mxk || act[lx + q][ly][lz];
                myk = myk || act[lx][ly + q][lz];
                mzk = mzk || act[lx][ly][lz + q];
            }
            const long i = (long(x) * d.ny + y) * d.nz + z;
            const bool zmin = z == 0;
            const double pin = zmin ? spatial[i / d.nz] * pin_scale : 0.0;
            const double v = cfl_cell_masked(U, i, N, c.S0[ids[i]], g, eps, zmin, pin, rf, mxk,
                                             myk, mzk);
            if (!(v == v))
                bad = true;
            else
                mx = fmax(mx, v);
        }
    }
    block_max_to(mx, bad, red, a.out + b);
}

// CFL maximum per box over all its cells.  Tiles outside the live set hold
// the exact zero state: their cells take the vacuum value of their material
// (and of the inflow on z-min cells) without reading the state.
// Every cell of one tile that may hold a non-zero state.
__device__ __forceinline__ void cfl_live_tile(const BatchCflArgs& a, const BoxDesc& d,
                                              const TileCoord& tc, int b, double* red)
{
    const StepCoef& c = a.coef[b];
    const GridDims g{d.nx, d.ny, d.nz, a.dx, a.dy, a.dz};
    const long N = long(d.nx) * d.ny * d.nz;
    const double eps = c.eig_eps;
    const double pin_scale = c.ed[0] * c.inv_alpha, rf = c.rf;
    const uint8_t* ids = a.ids + d.cell_off;
    const double* spatial = a.spatial + d.col_off;
    const int tid = threadIdx.x;
    const int x0 = tc.tx * TX, y0 = tc.ty * TY, z0 = tc.tz * TZ;
    const int x1 = min(x0 + TX, d.nx);
    double mx = 0.0;
    bool bad = false;
    const float* U = a.U + 4 * d.cell_off;
    const int z = z0 + tid % TZ, y = y0 + tid / TZ;
    if (z < d.nz && y < d.ny) {
        // unrolled over the tile's x planes so that the loads of all of a
        // lane's cells are in flight together
#pragma unroll
        for (int k = 0; k < TX; ++k) {
            const int x = x0 + k;
            if (x < x1) {
                const long i = (long(x) * d.ny + y) * d.nz + z;
                const bool zmin = z == 0;
                const double pin = zmin ? spatial[i / d.nz] * pin_scale : 0.0;
                // with the tail floor, a cell at or below the tile threshold
                // counts as the zero state it becomes when its tile is left out
                const double v =
                    a.tail_floor && !(U[i] > a.act_threshold)
                        ? cfl_vacuum(c.S0[ids[i]], g, a.wave_mode, eps, zmin, pin, rf)
                        : cfl_cell(U, i, N, c.S0[ids[i]], g, a.wave_mode, eps, zmin, pin, rf);
                if (!(v == v))
                    bad = true;
                else
                    mx = fmax(mx, v);
            }
        }
    }
    block_max_to(mx, bad, red, a.out + b);
}

// Live-list form of the plain CFL: the workgroups walk the compacted list of
// live tiles.
__global__ void __launch_bounds__(NT) cfl_list_kernel(BatchCflArgs a)
{
    __shared__ double red[NT / 64];
    const int n = *a.count;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int t = a.list[i];
        const int b = a.tile_box[t];
        if (a.step[b].active) {
            const BoxDesc d = a.box[b];
            cfl_live_tile(a, d, tile_coord(d, b, t), b, red);
        }
        __syncthreads();  // red is reused by the next tile
    }
}

// Live-list form, tiles outside the live set (exact zero state), one thread
// per tile: the tile's materials go to its box's mask, whose vacuum values
// the host adds.  A box's tiles are contiguous, so a wave almost always
// covers one box: its lanes' masks are OR-reduced and one lane updates the
// box (an atomic only when the mask still lacks a bit).
__global__ void __launch_bounds__(256) cfl_vacuum_tiles_kernel(BatchCflArgs a)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    int b = -1;
    unsigned long long m = 0;
    if (t < a.n_tiles) {
        b = a.tile_box[t];
        if (a.step[b].active && !a.live[t])
            m = a.tile_mats[t];
    }
    const int b0 = __shfl(b, 0, 64);
    if (__all(b == b0 || b < 0)) {
        for (int off = 32; off > 0; off >>= 1)
            m |= __shfl_xor(m, off, 64);
        if ((threadIdx.x & 63) == 0 && b0 >= 0 && m && (a.box_mats[b0] & m) != m)
            atomicOr(a.box_mats + b0, m);
    } else if (m && (a.box_mats[b] & m) != m) {
        atomicOr(a.box_mats + b, m);
    }
}

// Live-list form, z-min columns of the tiles outside the live set, one
// thread per column: the vacuum cell with the inflow speeds on z.
__global__ void __launch_bounds__(256) cfl_inflow_columns_kernel(BatchCflArgs a)
{
    __shared__ double red[256 / 64];
    const int b = blockIdx.y;
    if (!a.step[b].active)
        return;
    const BoxDesc d = a.box[b];
    const long col = long(blockIdx.x) * 256 + threadIdx.x;
    double mx = 0.0;
    if (col < long(d.nx) * d.ny) {
        const int x = int(col / d.ny), y = int(col % d.ny);
        const int t = d.tile_off + ((x / TX) * d.nty + y / TY) * d.ntz;
        if (!a.live[t]) {
            const StepCoef& c = a.coef[b];
            const GridDims g{d.nx, d.ny, d.nz, a.dx, a.dy, a.dz};
            const double pin = a.spatial[d.col_off + col] * (c.ed[0] * c.inv_alpha);
            mx = cfl_vacuum(c.S0[a.ids[d.cell_off + col * d.nz]], g, a.wave_mode, c.eig_eps, true,
                            pin, c.rf);
        }
    }
    block_max_to(mx, false, red, a.out + b);
}

__global__ void __launch_bounds__(NT) cfl_kernel(BatchCflArgs a)
{
    __shared__ double red[NT / 64];
    const int t = blockIdx.x;
    const int b = a.tile_box[t];
    if (!a.step[b].active)
        return;
    if (a.vacuum_aware) {
        if (!a.live[t])
            return;
        const BoxDesc d = a.box[b];
        cfl_tile_vacuum_aware(a, d, tile_coord(d, b, t), b, red);
        return;
    }
    const BoxDesc d = a.box[b];
    const TileCoord tc = tile_coord(d, b, t);
    if (!a.live || a.live[t]) {
        cfl_live_tile(a, d, tc, b, red);
        return;
    }
    const StepCoef& c = a.coef[b];
    const GridDims g{d.nx, d.ny, d.nz, a.dx, a.dy, a.dz};
    const double eps = c.eig_eps;
    const double pin_scale = c.ed[0] * c.inv_alpha, rf = c.rf;
    const uint8_t* ids = a.ids + d.cell_off;
    const double* spatial = a.spatial + d.col_off;
    const int tid = threadIdx.x;
    const int x0 = tc.tx * TX, y0 = tc.ty * TY;
    double mx = 0.0;
    const bool bad = false;
    {
        const unsigned mats = a.tile_mats[t];
        if (tid < MAXMAT && (mats >> tid & 1u))
            mx = cfl_vacuum(c.S0[tid], g, a.wave_mode, eps, false, 0.0, rf);
        if (tc.tz == 0 && tid < TX * TY) {
            const int x = x0 + tid / TY, y = y0 + tid % TY;
            if (x < d.nx && y < d.ny) {
                const long col = long(x) * d.ny + y;
                const double pin = spatial[col] * pin_scale;
                mx = fmax(mx, cfl_vacuum(c.S0[ids[col * d.nz]], g, a.wave_mode, eps, true, pin,
                                         rf));
            }
        }
    }
    block_max_to(mx, bad, red, a.out + b);
}

__global__ void boundary_kernel(BatchBoundaryArgs a)
{
    const int b = blockIdx.y;
    if (!a.step[b].active)
        return;
    const BoxDesc d = a.box[b];
    const long m2 = long(d.nx) * d.ny;
    const long col = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (col >= m2)
        return;
    boundary_column(a.coef[b], a.ids[d.cell_off + col * d.nz], a.spatial[d.col_off + col], col, m2,
                    a.integrating_factor != 0, a.bnd + 12 * d.col_off);
}

// Next-step tile set of every active box: the single-box rule (a tile is
// evaluated when it holds phi above the threshold, when a cell of a
// neighbouring tile within TILE_REACH cells of it does, or when it is a z = 0
// tile while the inflow is on) per box.  The tile-skip network may leave out,
// for the steps its countdown runs, tiles the reach rule adds while no cell
// within TILE_GUARD cells of them holds phi above the threshold (checked
// every step).  Also records, when asked, those candidate tiles (ring) and
// the running max of tile_max.
__global__ void mask_kernel(BatchMaskArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_tiles)
        return;
    const int b = a.tile_box[t];
    const BoxStep st = a.step[b];
    if (!st.active || (a.only && !a.only[b])) {
        a.next[t] = 0;
        if (a.ring)
            a.ring[t] = 0;
        return;
    }
    const BoxDesc d = a.box[b];
    const TileCoord tc = tile_coord(d, b, t);
    const float self_max = a.computed[t] ? a.tile_max[t] : 0.f;
    if (a.wmax)
        a.wmax[t] = fmaxf(a.wmax[t], self_max);
    const bool forced = st.inflow_on && tc.tz == 0;
    const bool self_on = self_max > a.threshold;
    bool on = forced || self_on;
    for (int dx = -a.rx; dx <= a.rx && !on; ++dx) {
        const int xx = tc.tx + dx;
        if (xx < 0 || xx >= d.ntx)
            continue;
        for (int dy = -a.ry; dy <= a.ry && !on; ++dy) {
            const int yy = tc.ty + dy;
            if (yy < 0 || yy >= d.nty)
                continue;
            for (int dz = -a.rz; dz <= a.rz && !on; ++dz) {
                const int zz = tc.tz + dz;
                if (zz < 0 || zz >= d.ntz || (dx == 0 && dy == 0 && dz == 0))
                    continue;
                const int u = d.tile_off + (xx * d.nty + yy) * d.ntz + zz;
                on = a.computed[u] && facing_bound(a.tile_face + FACE_SLOTS * long(u), dx, dy, dz,
                                                   0) > a.threshold;
            }
        }
    }
    // Candidates of the network: added by the reach rule alone and, with the
    // guard, with nothing above the threshold within TILE_GUARD cells.
    bool cand = false;
    uint8_t ring = 0;
    if (on && !forced && !self_on && (a.ring || a.pred)) {
        bool near = false;
        for (int dx = -1; dx <= 1 && !near; ++dx) {
            const int xx = tc.tx + dx;
            if (xx < 0 || xx >= d.ntx)
                continue;
            for (int dy = -1; dy <= 1 && !near; ++dy) {
                const int yy = tc.ty + dy;
                if (yy < 0 || yy >= d.nty)
                    continue;
                for (int dz = -1; dz <= 1 && !near; ++dz) {
                    const int zz = tc.tz + dz;
                    if (zz < 0 || zz >= d.ntz || (dx == 0 && dy == 0 && dz == 0))
                        continue;
                    const int u = d.tile_off + (xx * d.nty + yy) * d.ntz + zz;
                    near = a.computed[u] && facing_bound(a.tile_face + FACE_SLOTS * long(u), dx,
                                                         dy, dz, 6) > a.threshold;
                }
            }
        }
        cand = a.guard ? !near : true;
        ring = cand ? (near ? 2 : 1) : 0;
    }
    if (a.ring)
        a.ring[t] = ring;
    if (a.pred && a.pred[t] > 0) {
        if (cand && st.model_on)
            on = false;
        a.pred[t] -= 1;
    }
    a.next[t] = (a.all_on || st.full_set || on) ? 1 : 0;
}

// Features of every tile of the active boxes and, with a model, the
// network's prediction for the tiles evaluated in the last step.  One wave
// per 64 tiles; the network's activations live in LDS (m1_mlp.h).
__global__ void __launch_bounds__(mlp::WG) feature_kernel(BatchFeatureArgs a)
{
    if (a.gate && !*a.gate)
        return;  // uniform: no barrier has run
    __shared__ uint16_t wl[mlp::NPARAM];
    __shared__ float A[mlp::NH][mlp::WG], B[mlp::NH][mlp::WG];
    if (a.w16) {
        for (int k = threadIdx.x; k < mlp::NPARAM; k += blockDim.x)
            wl[k] = a.w16[k];
        __syncthreads();
    }
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_tiles)
        return;  // no barrier follows
    const int b = a.tile_box[t];
    const BoxStep st = a.step[b];
    const BoxDesc d = a.box[b];
    const TileCoord tc = tile_coord(d, b, t);
    auto lg = [](float v) { return log10f(fmaxf(v, 1e-30f)) * 0.1f; };
    auto tmax = [&](int u) { return a.computed[u] ? a.tile_max[u] : 0.f; };
    const float mx = tmax(t);
    const float ps = a.computed[t] ? a.tile_sums[2 * long(t)] : 0.f;
    const float js = a.computed[t] ? a.tile_sums[2 * long(t) + 1] : 0.f;
    float x[mlp::NIN];
    x[0] = lg(mx);
    x[1] = lg(ps * (1.f / float(TX * TY * TZ)));
    x[2] = ps > 0.f ? fminf(js / ps, 1.f) : 0.f;
    x[3] = a.tile_rho[t] * 1e-3f;
    x[4] = st.energy * (1.f / 250.f);
    const int nb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        // the face neighbour's slab within TILE_REACH cells of this tile
        const int xx = tc.tx + nb[k][0], yy = tc.ty + nb[k][1], zz = tc.tz + nb[k][2];
        const bool in = xx >= 0 && xx < d.ntx && yy >= 0 && yy < d.nty && zz >= 0 && zz < d.ntz;
        const int u = in ? d.tile_off + (xx * d.nty + yy) * d.ntz + zz : 0;
        const float v = in && a.computed[u]
                            ? facing_bound(a.tile_face + FACE_SLOTS * long(u), nb[k][0], nb[k][1],
                                           nb[k][2], 0)
                            : 0.f;
        x[5 + k] = lg(v);
    }
    if (a.features && st.active)
#pragma unroll
        for (int k = 0; k < mlp::NIN; ++k)
            a.features[long(t) * mlp::NIN + k] = x[k];
    const bool forced = st.inflow_on && tc.tz == 0;
    if (a.w16 && a.computed[t] && st.active && st.model_on) {
        bool drop = false;
        if (!forced && !(mx > a.threshold)) {
            const float z = mlp::forward_one(wl, a.params, x, A, B, threadIdx.x, nullptr,
                                             nullptr, nullptr);
            drop = mlp::sigmoid(z) > a.decision;
        }
        a.pred[t] = drop ? uint8_t(a.horizon) : uint8_t(0);
    }
}

// Recording: one sample per ring tile saved `horizon` steps ago.
__global__ void label_kernel(BatchLabelArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_tiles || !a.cand[t])
        return;
    const int k = atomicAdd(a.count, 1);
    if (k >= a.cap)
        return;
    for (int i = 0; i < mlp::NIN; ++i)
        a.out_x[long(k) * mlp::NIN + i] = a.features[long(t) * mlp::NIN + i];
    // label (1: stayed at or below the threshold), plus 2 when a cell within
    // TILE_GUARD cells was above it when the sample was taken
    a.out_y[k] = (a.wmax[t] <= a.threshold ? 1.f : 0.f) + (a.cand[t] == 2 ? 2.f : 0.f);
}

// Judge: largest |full - skip| and |full| dose increment of each box.
__global__ void compare_kernel(BatchCompareArgs a)
{
    __shared__ float red[2][256 / 64];
    const int b = blockIdx.y;
    if (!a.judge[b])
        return;
    const BoxDesc d = a.box[b];
    const long n = long(d.nx) * d.ny * d.nz;
    float df = 0.f, fl = 0.f;
    for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += long(gridDim.x) * blockDim.x) {
        const long c = d.cell_off + i;
        const float f = a.full_hi[c] + a.full_lo[c];
        const float k = a.skip_hi[c] + a.skip_lo[c];
        df = fmaxf(df, fabsf(f - k));
        fl = fmaxf(fl, fabsf(f));
    }
    for (int off = 32; off > 0; off >>= 1) {
        df = fmaxf(df, __shfl_xor(df, off, 64));
        fl = fmaxf(fl, __shfl_xor(fl, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = df;
        red[1][threadIdx.x >> 6] = fl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < int(blockDim.x / 64); ++w) {
            df = fmaxf(df, red[0][w]);
            fl = fmaxf(fl, red[1][w]);
        }
        df = fmaxf(df, red[0][0]);
        fl = fmaxf(fl, red[1][0]);
        atomicMax(a.max_diff + b, __float_as_uint(df));
        atomicMax(a.max_full + b, __float_as_uint(fl));
    }
}

__device__ __forceinline__ void slice_cell(const BatchSliceArgs& a, const BoxDesc& d, int sel,
                                           long n, long i);

// One grid row per box; the ~4 judged boxes of a step are copied by the
// workgroups of their rows, the others' rows exit at once.
__global__ void slices_kernel(BatchSliceArgs a)
{
    const int b = blockIdx.y;
    const int sel = a.sel[b];
    if (!sel)
        return;
    const BoxDesc d = a.box[b];
    const long n = long(d.nx) * d.ny * d.nz;
    for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += long(gridDim.x) * blockDim.x)
        slice_cell(a, d, sel, n, i);
}

__device__ __forceinline__ void slice_cell(const BatchSliceArgs& a, const BoxDesc& d, int sel,
                                           long n, long i)
{
    const long c = d.cell_off + i;
    switch (a.op) {
    case SLICE_SAVE: {
        const int k = sel - 1;
        for (int m = 0; m < 4; ++m)
            a.ck_U[k][4 * d.cell_off + m * n + i] = a.U[4 * d.cell_off + m * n + i];
        a.ck_hi[k][c] = a.dose_hi[c];
        a.ck_lo[k][c] = a.dose_lo[c];
        break;
    }
    case SLICE_RESTORE: {
        if (sel == 3) {
            for (int m = 0; m < 4; ++m)
                a.U[4 * d.cell_off + m * n + i] = 0.f;
            a.dose_hi[c] = 0.f;
            a.dose_lo[c] = 0.f;
        } else {
            const int k = sel - 1;
            for (int m = 0; m < 4; ++m)
                a.U[4 * d.cell_off + m * n + i] = a.ck_U[k][4 * d.cell_off + m * n + i];
            a.dose_hi[c] = a.ck_hi[k][c];
            a.dose_lo[c] = a.ck_lo[k][c];
        }
        break;
    }
    case SLICE_RESET:
        a.full_hi[c] = a.full_lo[c] = a.skip_hi[c] = a.skip_lo[c] = 0.f;
        break;
    default: {
        float hi = a.dose_hi[c], lo = a.dose_lo[c];
        two_sum_add(hi, lo, a.skip_hi[c]);
        a.dose_hi[c] = hi;
        a.dose_lo[c] = lo + a.skip_lo[c];
    }
    }
}

// One workgroup: ascending list of the set flags (each thread scans one
// contiguous chunk; chunk counts are prefix-summed in LDS).  Used when the
// chunk counts of the multi-workgroup form would not fit after the header.
constexpr int COMPACT_THREADS = 1024;
constexpr int COMPACT_CHUNK = 1024;

__global__ void __launch_bounds__(COMPACT_THREADS)
    compact_kernel(const uint8_t* flags, int n, int* list, int* count)
{
    if (threadIdx.x < 3 * XCD_QUEUES)
        count[1 + threadIdx.x] = 0;  // the stage queues of this list
    __shared__ int sums[COMPACT_THREADS];
    const int tid = threadIdx.x;
    const int per = (n + COMPACT_THREADS - 1) / COMPACT_THREADS;
    const int lo = min(tid * per, n), hi = min(lo + per, n);
    int c = 0;
    for (int i = lo; i < hi; ++i)
        c += flags[i] != 0;
    sums[tid] = c;
    __syncthreads();
    for (int off = 1; off < COMPACT_THREADS; off <<= 1) {
        const int v = tid >= off ? sums[tid - off] : 0;
        __syncthreads();
        sums[tid] += v;
        __syncthreads();
    }
    int pos = sums[tid] - c;
    for (int i = lo; i < hi; ++i)
        if (flags[i])
            list[pos++] = i;
    if (tid == COMPACT_THREADS - 1)
        *count = sums[tid];
}

// Multi-workgroup form: workgroup g takes flags [g C, (g + 1) C), C = 1024,
// four consecutive flags per thread.  Pass 1 writes each chunk's count after
// the list header; pass 2 adds the counts of the chunks before its own and
// writes its entries in order, so the list is the same ascending one.
__device__ __forceinline__ int block_exclusive_scan(int v, int* wsum, int& total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off)
            x += y;
    }
    if (lane == 63)
        wsum[w] = x;
    __syncthreads();
    int before = 0;
    total = 0;
    for (int k = 0; k < 256 / 64; ++k) {
        if (k < w)
            before += wsum[k];
        total += wsum[k];
    }
    return before + x - v;
}

__device__ __forceinline__ int chunk_flags(const uint8_t* flags, int n, int i0, bool f[4])
{
    int c = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        f[k] = i0 + k < n && flags[i0 + k] != 0;
        c += f[k];
    }
    return c;
}

__global__ void __launch_bounds__(256)
    compact_count_kernel(const uint8_t* flags, int n, int* count)
{
    __shared__ int wsum[256 / 64];
    if (blockIdx.x == 0 && threadIdx.x < 3 * XCD_QUEUES)
        count[1 + threadIdx.x] = 0;  // the stage queues of this list
    bool f[4];
    const int c = chunk_flags(flags, n, blockIdx.x * COMPACT_CHUNK + 4 * threadIdx.x, f);
    int total;
    block_exclusive_scan(c, wsum, total);
    if (threadIdx.x == 0)
        count[LIST_HEADER + blockIdx.x] = total;
}

__global__ void __launch_bounds__(256)
    compact_write_kernel(const uint8_t* flags, int n, int* list, int* count)
{
    __shared__ int wsum[256 / 64];
    __shared__ int s_off;
    const int* chunk = count + LIST_HEADER;
    int off = 0;
    for (int j = threadIdx.x; j < int(blockIdx.x); j += 256)
        off += chunk[j];
    for (int d = 32; d > 0; d >>= 1)
        off += __shfl_xor(off, d, 64);
    if ((threadIdx.x & 63) == 0)
        wsum[threadIdx.x >> 6] = off;
    __syncthreads();
    if (threadIdx.x == 0)
        s_off = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    const int base = s_off;
    __syncthreads();  // wsum is reused by the scan
    bool f[4];
    const int i0 = blockIdx.x * COMPACT_CHUNK + 4 * threadIdx.x;
    const int c = chunk_flags(flags, n, i0, f);
    int total;
    int pos = base + block_exclusive_scan(c, wsum, total);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (f[k])
            list[pos++] = i0 + k;
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
        *count = base + total;
}

// Tiles that leave the evaluated set become exact vacuum (U = V1 = V2 = 0).
__device__ __forceinline__ void zero_tile(const BatchZeroArgs& a, int t);

// Workgroup g checks tiles [g NT, (g + 1) NT) at once (one lane each), then
// zeroes the leaving ones among them.  (A capped grid walking the tiles one
// flag load after another was latency-bound: ~150 us per step on the
// 256-beamlet batch.)
__global__ void __launch_bounds__(NT) zero_kernel(BatchZeroArgs a)
{
    __shared__ int s_list[NT];
    __shared__ int s_n;
    if (threadIdx.x == 0)
        s_n = 0;
    __syncthreads();
    const int t = int(blockIdx.x) * NT + int(threadIdx.x);
    if (t < a.n_tiles && a.computed[t] && !a.next[t])
        s_list[atomicAdd(&s_n, 1)] = t;
    __syncthreads();
    const int n = s_n;
    for (int k = 0; k < n; ++k)
        zero_tile(a, s_list[k]);
}

__device__ __forceinline__ void zero_tile(const BatchZeroArgs& a, int t)
{
    const int b = a.tile_box[t];
    const BoxDesc d = a.box[b];
    const TileCoord tc = tile_coord(d, b, t);
    const long N = long(d.nx) * d.ny * d.nz;
    const int z = tc.tz * TZ + threadIdx.x % TZ;
    const int y = tc.ty * TY + threadIdx.x / TZ;
    if (z >= d.nz || y >= d.ny)
        return;
    const long off4 = 4 * d.cell_off;
    for (int x = tc.tx * TX; x < min(tc.tx * TX + TX, d.nx); ++x) {
        const long i = (long(x) * d.ny + y) * d.nz + z;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            a.U[off4 + i + c * N] = 0.f;
            a.V1[off4 + i + c * N] = 0.f;
            a.V2[off4 + i + c * N] = 0.f;
            if (a.W1) {
                a.W1[off4 + i + c * N] = 0.f;
                a.W2[off4 + i + c * N] = 0.f;
            }
            if (a.R1) {
                a.R1[off4 + i + c * N] = 0.f;
                a.R2[off4 + i + c * N] = 0.f;
            }
        }
    }
}

__global__ void scatter_dose_kernel(const float* hi, const float* lo, int nx, int ny, int nz,
                                    int x0, int y0, int py, int pz, double* phantom)
{
    const long n = long(nx) * ny * nz;
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int z = int(i % nz);
    const int y = int((i / nz) % ny);
    const int x = int(i / (long(nz) * ny));
    phantom[(long(x0 + x) * py + (y0 + y)) * pz + z] += double(hi[i]) + double(lo[i]);
}

__global__ void box_dose_kernel(const float* hi, const float* lo, long n, double* out)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = double(hi[i]) + double(lo[i]);
}

}  // namespace

int stage_blocks_per_cu()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, reinterpret_cast<const void*>(&stage_kernel<STAGE2, true, false, LAUNCH_TILE>),
            NT, 0) !=
            hipSuccess ||
        n < 1)
        n = 1;
    return n;
}

template <int KIND, bool IFAC, bool D>
int stage_form_of(const void* func)
{
    if (func == reinterpret_cast<const void*>(&stage_kernel<KIND, IFAC, D, LAUNCH_TILE>))
        return LAUNCH_TILE;
    if (func == reinterpret_cast<const void*>(&stage_kernel<KIND, IFAC, D, LAUNCH_OVERFLOW>))
        return LAUNCH_OVERFLOW;
    if (func == reinterpret_cast<const void*>(&stage_kernel<KIND, IFAC, D, LAUNCH_QUEUE>))
        return LAUNCH_QUEUE;
    return -1;
}

int stage_form(const void* func)
{
    int f = -1;
    for (int d = 0; d < 2 && f < 0; ++d) {
        const bool dd = d != 0;
        f = dd ? stage_form_of<STAGE0, false, true>(func) : stage_form_of<STAGE0, false, false>(func);
        if (f < 0)
            f = dd ? stage_form_of<STAGE1, false, true>(func) : stage_form_of<STAGE1, false, false>(func);
        if (f < 0)
            f = dd ? stage_form_of<STAGE1, true, true>(func) : stage_form_of<STAGE1, true, false>(func);
        if (f < 0)
            f = dd ? stage_form_of<STAGE2, false, true>(func) : stage_form_of<STAGE2, false, false>(func);
        if (f < 0)
            f = dd ? stage_form_of<STAGE2, true, true>(func) : stage_form_of<STAGE2, true, false>(func);
    }
    return f;
}

hipError_t launch_stage(int kind, bool ifac, const BatchStageArgs& a, hipStream_t s)
{
    // persistent (queue set): n_list is the grid (workgroups resident at
    // once); estimated (strided > 0): strided one-tile workgroups, then
    // n_list - strided overflow workgroups; otherwise n_list workgroups
    if (a.n_list <= 0)
        return hipSuccess;
    const int form = a.queue ? LAUNCH_QUEUE : LAUNCH_TILE;
    const unsigned main_grid = unsigned(a.queue || !a.strided ? a.n_list : a.strided);
    const unsigned over_grid = unsigned(a.queue || !a.strided ? 0 : a.n_list - a.strided);
    const dim3 block(NT);
#define M1_LAUNCH_FORM(K, I, DD, L, G)                                                    \
    hipLaunchKernelGGL((stage_kernel<K, I, DD, L>), dim3(G), block, 0, s, a)
#define M1_LAUNCH(K, I)                                                                   \
    do {                                                                                  \
        if (a.f64_closure) {                                                              \
            if (form == LAUNCH_QUEUE)                                                     \
                M1_LAUNCH_FORM(K, I, true, LAUNCH_QUEUE, main_grid);                      \
            else                                                                          \
                M1_LAUNCH_FORM(K, I, true, LAUNCH_TILE, main_grid);                       \
            if (over_grid)                                                                \
                M1_LAUNCH_FORM(K, I, true, LAUNCH_OVERFLOW, over_grid);                   \
        } else {                                                                          \
            if (form == LAUNCH_QUEUE)                                                     \
                M1_LAUNCH_FORM(K, I, false, LAUNCH_QUEUE, main_grid);                     \
            else                                                                          \
                M1_LAUNCH_FORM(K, I, false, LAUNCH_TILE, main_grid);                      \
            if (over_grid)                                                                \
                M1_LAUNCH_FORM(K, I, false, LAUNCH_OVERFLOW, over_grid);                  \
        }                                                                                 \
    } while (0)
    switch (kind) {
    case STAGE0:
        M1_LAUNCH(STAGE0, false);
        break;
    case STAGE1:
        if (ifac)
            M1_LAUNCH(STAGE1, true);
        else
            M1_LAUNCH(STAGE1, false);
        break;
    case STAGE2:
        if (ifac)
            M1_LAUNCH(STAGE2, true);
        else
            M1_LAUNCH(STAGE2, false);
        break;
    case STAGE_RELAX:
        M1_LAUNCH(STAGE_RELAX, false);
        break;
    default:
        return hipErrorInvalidValue;
    }
#undef M1_LAUNCH
#undef M1_LAUNCH_FORM
    return hipGetLastError();
}

hipError_t launch_cfl(const BatchCflArgs& a, hipStream_t s)
{
    if (a.n_tiles <= 0)
        return hipSuccess;
    if (a.list && a.live && !a.vacuum_aware) {
        hipLaunchKernelGGL(cfl_vacuum_tiles_kernel, dim3(unsigned((a.n_tiles + 255) / 256)),
                           dim3(256), 0, s, a);
        if (a.n_boxes > 0 && a.max_columns > 0)
            hipLaunchKernelGGL(cfl_inflow_columns_kernel,
                               dim3(unsigned((a.max_columns + 255) / 256), unsigned(a.n_boxes)),
                               dim3(256), 0, s, a);
        // a grid of resident workgroups walking the live list (per device)
        static std::atomic<int> grids[64];
        int dev = 0;
        (void)hipGetDevice(&dev);
        int grid = dev >= 0 && dev < 64 ? grids[dev].load() : 0;
        if (grid == 0) {
            int cus = 0, per = 0;
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, cfl_list_kernel, NT, 0);
            grid = (cus > 0 ? cus : 256) * (per > 0 ? per : 4);
            if (dev >= 0 && dev < 64)
                grids[dev].store(grid);
        }
        hipLaunchKernelGGL(cfl_list_kernel, dim3(unsigned(min(grid, a.n_tiles))), dim3(NT), 0,
                           s, a);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(cfl_kernel, dim3(unsigned(a.n_tiles)), dim3(NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_boundary(const BatchBoundaryArgs& a, int max_columns, hipStream_t s)
{
    if (a.n_boxes <= 0 || max_columns <= 0)
        return hipSuccess;
    const dim3 grid(unsigned((max_columns + 255) / 256), unsigned(a.n_boxes));
    hipLaunchKernelGGL(boundary_kernel, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mask(const BatchMaskArgs& a, hipStream_t s)
{
    if (a.n_tiles <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(mask_kernel, dim3(unsigned((a.n_tiles + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_compact(const uint8_t* flags, int n, int* list, int* count, hipStream_t s)
{
    const int blocks = (n + COMPACT_CHUNK - 1) / COMPACT_CHUNK;
    if (blocks >= 1 && blocks <= COMPACT_MAX_BLOCKS) {
        hipLaunchKernelGGL(compact_count_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, flags,
                           n, count);
        hipLaunchKernelGGL(compact_write_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, flags,
                           n, list, count);
    } else {
        hipLaunchKernelGGL(compact_kernel, dim3(1), dim3(COMPACT_THREADS), 0, s, flags, n, list,
                           count);
    }
    return hipGetLastError();
}

hipError_t launch_zero(const BatchZeroArgs& a, hipStream_t s)
{
    if (a.n_tiles <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(zero_kernel, dim3(unsigned((a.n_tiles + NT - 1) / NT)), dim3(NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_scatter_dose(const float* hi, const float* lo, int nx, int ny, int nz, int x0,
                               int y0, int py, int pz, double* phantom, hipStream_t s)
{
    const long n = long(nx) * ny * nz;
    hipLaunchKernelGGL(scatter_dose_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, hi,
                       lo, nx, ny, nz, x0, y0, py, pz, phantom);
    return hipGetLastError();
}

hipError_t launch_features(const BatchFeatureArgs& a, hipStream_t s)
{
    if (a.n_tiles <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(feature_kernel, dim3(unsigned((a.n_tiles + mlp::WG - 1) / mlp::WG)),
                       dim3(mlp::WG), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_labels(const BatchLabelArgs& a, hipStream_t s)
{
    if (a.n_tiles <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(label_kernel, dim3(unsigned((a.n_tiles + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_compare(const BatchCompareArgs& a, int max_cells, hipStream_t s)
{
    if (a.n_boxes <= 0 || max_cells <= 0)
        return hipSuccess;
    const unsigned gx = unsigned(std::min((max_cells + 255) / 256, 64));
    hipLaunchKernelGGL(compare_kernel, dim3(gx, unsigned(a.n_boxes)), dim3(256), 0, s, a);
    return hipGetLastError();
}

__global__ void clear_kernel(unsigned long long* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0ull;
}

hipError_t launch_clear(unsigned long long* p, int n, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(clear_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_slices(const BatchSliceArgs& a, int max_cells, hipStream_t s)
{
    if (a.n_boxes <= 0 || max_cells <= 0)
        return hipSuccess;
    const unsigned gx = unsigned(std::min((max_cells + 255) / 256, 64));
    hipLaunchKernelGGL(slices_kernel, dim3(gx, unsigned(a.n_boxes)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Live tiles of active boxes holding a quasi-steady material.
__global__ void air_mask_kernel(BatchAirArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_tiles)
        return;
    a.air[t] = a.live[t] && a.step[a.tile_box[t]].active && (a.tile_mats[t] & a.qs_mask) ? 1 : 0;
}

// The listed tiles' quasi-steady cells take the relaxed state (from -> to).
__global__ void __launch_bounds__(NT) air_commit_kernel(BatchAirArgs a)
{
    const int n = *a.count;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const int t = a.list[k];
        const int b = a.tile_box[t];
        const BoxDesc d = a.box[b];
        const TileCoord tc = tile_coord(d, b, t);
        const long N = long(d.nx) * d.ny * d.nz;
        const int z = tc.tz * TZ + threadIdx.x % TZ;
        const int y = tc.ty * TY + threadIdx.x / TZ;
        if (z >= d.nz || y >= d.ny)
            continue;
        const long off4 = 4 * d.cell_off;
        for (int x = tc.tx * TX; x < min(tc.tx * TX + TX, d.nx); ++x) {
            const long i = (long(x) * d.ny + y) * d.nz + z;
            if (!(a.qs_mask >> a.ids[d.cell_off + i] & 1u))
                continue;
#pragma unroll
            for (int c = 0; c < 4; ++c)
                a.to[off4 + i + c * N] = a.from[off4 + i + c * N];
        }
    }
}

__global__ void air_reset_kernel(BatchAirArgs a)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.n_boxes)
        return;
    a.done[b] = 0;
    a.change[b] = 0u;
    a.change[2 * a.n_boxes + b] = 0x7f800000u;  // +inf: the first sweep runs
}

hipError_t launch_air_mask(const BatchAirArgs& a, hipStream_t s)
{
    if (a.n_tiles <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(air_mask_kernel, dim3(unsigned((a.n_tiles + 255) / 256)), dim3(256), 0, s,
                       a);
    return hipGetLastError();
}

hipError_t launch_air_reset(const BatchAirArgs& a, hipStream_t s)
{
    if (a.n_boxes <= 0 || !a.done)
        return hipSuccess;
    hipLaunchKernelGGL(air_reset_kernel, dim3(unsigned((a.n_boxes + 255) / 256)), dim3(256), 0, s,
                       a);
    return hipGetLastError();
}

hipError_t launch_air_commit(const BatchAirArgs& a, int grid, hipStream_t s)
{
    if (grid <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(air_commit_kernel, dim3(unsigned(grid)), dim3(NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_box_dose(const float* hi, const float* lo, long n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(box_dose_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, hi, lo,
                       n, out);
    return hipGetLastError();
}

}  // namespace M1_BATCH_NS

const BatchKernels& M1_BATCH_TABLE()
{
    using namespace M1_BATCH_NS;
    static const BatchKernels k{F32_TX,         F32_TY,          F32_TZ,
                                &launch_stage,  &launch_cfl,     &launch_boundary,
                                &launch_mask,   &launch_compact, &launch_zero,
                                &launch_scatter_dose, &launch_box_dose, &launch_features,
                                &launch_labels, &launch_compare, &launch_slices,
                                &launch_clear,  &stage_blocks_per_cu, &stage_form,
                                &launch_air_mask, &launch_air_commit, &launch_air_reset};
    return k;
}
}  // namespace m1
