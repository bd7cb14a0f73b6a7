// Device-side data of the MI355X march: per-step coefficients, launch
// arguments and the launchers of the kernels in m1_kernels.hip.
//
// m1_kernels.hip is compiled twice: with M1_EXACT=1 (namespace m1::exact,
// every multiply and add rounded separately and every division kept, the
// operation order of the reference's NumPy/Numba arithmetic) and with
// M1_EXACT=0 (namespace m1::fast, fused multiply-adds and reciprocal
// constants).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace m1 {

constexpr int MAXMAT = 8;
constexpr double MEV_TO_JOULE = 1.602176634e-13;  // numerics/energy.py:37

// Scalars of one energy step, per material m (index = material id).
// Energies: E (step start), E1 = E - dE, E2 = E - dE/2, E13 = E - dE/3,
// E23 = E - 2 dE/3.  C(tau) = diag(exp(-min(tau_n,700)), exp(-min(tau_n+tau_s,700)) x3)
// is the integrating factor of energy.py:273-310 ("p" = phi entry, "j" = J entries).
struct StepCoef {
    double dE;
    double ed[4];            // inflow spectrum N(E) at E, E13, E23, E1 (boundary.py:90-99)
    double rf;               // exp(-theta^2), boundary.py:102
    double S0[MAXMAT];       // S(E)
    double S1[MAXMAT];       // S(E1)
    double S2[MAXMAT];       // S(E2)
    double S13[MAXMAT];      // S(E13)
    double S23[MAXMAT];      // S(E23)
    double f1p[MAXMAT], f1j[MAXMAT];    // C(tau(dE))
    double f2p[MAXMAT], f2j[MAXMAT];    // C(tau(dE/2))
    double i1p[MAXMAT], i1j[MAXMAT];    // C(tau(dE))^-1
    double i2p[MAXMAT], i2j[MAXMAT];    // C(tau(dE/2))^-1
    double i13p[MAXMAT], i13j[MAXMAT];  // C(tau(dE/3))^-1
    double i23p[MAXMAT], i23j[MAXMAT];  // C(tau(2dE/3))^-1
    double rho[MAXMAT];                 // density, kg/m^3
    double hn[MAXMAT], hs[MAXMAT];      // Strang half-step factors (energy.py:1168-1191)
    // Normalisation of the FP32-bulk march: the device state is U / alpha.
    double alpha;            // peak inflow fluence of the march
    double inv_alpha;
    double eig_eps;          // 1e-10 / alpha (wave_speeds.py:10 in scaled units)
    double weno_eps;         // 1e-12 / alpha^2 (weno_z.py:8 in scaled units)
};

// Float copy of the per-cell stage scalars for the FP32-bulk kernels.
struct StepCoef32 {
    float dE;
    float weno_eps;
    float S0[MAXMAT];
    float g1p[MAXMAT], g1j[MAXMAT];     // C(tau(dE)) / S(E1): V1 -> U1
    float g2p[MAXMAT], g2j[MAXMAT];     // C(tau(dE/2)) / S(E2): V2 -> U2
    float i1p[MAXMAT], i1j[MAXMAT];
    float i2p[MAXMAT], i2j[MAXMAT];
    float n1p[MAXMAT], n1j[MAXMAT];     // C(tau(dE)) / S(E1): V_new -> U_new
    float S1[MAXMAT], S2[MAXMAT];
    float dose_scale[MAXMAT];           // alpha * MEV / rho
    // step that multiplies the flux divergence: dE, or 0 for a quasi-steady
    // material (held through the step and relaxed after it); in the relaxation
    // coefficients, the pseudo-time CFL number of the quasi-steady materials
    float dEm[MAXMAT];
    double eig_eps;
};

struct GridDims {
    int nx, ny, nz;
    double dx, dy, dz;
};

enum StageKind : int {
    STAGE0 = 0,    // V1 = rhs(U) dE + S0 U
    STAGE1 = 1,    // V2 = 3/4 S0 U + 1/4 (C1^-1 rhs(U1) dE + V1)
    STAGE2 = 2,    // U  = proj(C1 (1/3 S0 U + 2/3 (C2^-1 rhs(U2) dE + V2)) / S1), dose
    STAGE_RHS = 3, // out = rhs(U)  (spatial_operator)
    STAGE_FE = 4,  // out = proj((rhs(U) dE + S0 U) / S1), edep = dE U0 S0
    STAGE_RELAX = 5  // out = proj(U + dtau rhs(U)) on quasi-steady cells (batch)
};

struct StageArgs {
    GridDims g;
    const double* src;      // stencil input: U (STAGE0/RHS/FE), V1 (STAGE1), V2 (STAGE2)
    double* U;              // accepted state U^n; STAGE2 overwrites it with U^{n+1}
    const double* V1;       // STAGE1: raw V1; STAGE2: raw V1 (phi only, for q1)
    double* out;            // STAGE0: V1, STAGE1: V2, RHS: rhs, FE: U^{n+1}
    double* dose;           // STAGE2 / FE: dose += edep * MEV / rho (nullable)
    double* edep;           // STAGE2 / FE: per-cell energy deposition (nullable)
    const uint8_t* ids;     // material ids
    const double* bnd;      // (4, nx, ny) z-min boundary state of this stage
    const StepCoef* coef;   // device copy of the step scalars
    int wave_mode;          // 0 exact M1 speeds, 1 unit speeds
    // Split form (all four set): the stencil input of stages 1 and 2 formed
    // once per cell into win (4 x N) and the face fluxes of each axis in
    // fx (4 x (nx+1) ny nz), fy (4 x nx (ny+1) nz), fz (4 x nx ny (nz+1));
    // otherwise the fused tile kernel.
    double* win;
    double* fx;
    double* fy;
    double* fz;
};

struct CflArgs {
    GridDims g;
    const double* U;
    const uint8_t* ids;
    const double* spatial;
    const StepCoef* coef;   // uses S0, ed[0], rf
    unsigned long long* out;  // max of the inverse step, as ordered bits
    int wave_mode;
    // Vacuum-aware CFL (energy.py:459-537): per-cell activity (nullable).  An
    // axis's speed counts only where an active cell lies within act_radius
    // cells along that axis.
    const uint8_t* act;
    int act_radius;
};

struct BoundaryArgs {
    GridDims g;
    const double* spatial;
    const uint8_t* ids;
    const StepCoef* coef;
    double* bnd;            // 3 x (4, nx, ny): stage 0 (U_in(E)), stage 1, stage 2
    int integrating_factor;
};

struct ScaleArgs {
    long n;
    double* U;
    const uint8_t* ids;
    const StepCoef* coef;
    int nuclear, scattering;
};

struct StageArgs32 {
    GridDims g;
    const float* src;       // stencil input: U (STAGE0/RHS/FE), V1 (STAGE1), V2 (STAGE2)
    float* U;               // accepted state (scaled); STAGE2 writes U^{n+1}
    const float* V1;
    float* out;             // STAGE0: V1, STAGE1: V2, RHS: rhs, FE: U^{n+1}
    float* dose_hi;         // compensated dose pair, Gy (nullable)
    float* dose_lo;
    float* edep;            // scaled energy deposition (nullable)
    const uint8_t* ids;
    const float* bnd;       // (4, nx, ny) scaled boundary state of this stage
    const StepCoef32* coef;
    int wave_mode;
    // Tile skipping (nullable): compute[t] != 0 for workgroup tiles to
    // evaluate; STAGE2 writes the tile's max phi to tile_max[t].
    const uint8_t* compute;
    float* tile_max;
    float* tile_face;       // FACE_SLOTS per tile (nullable; needs tile_max)
    int f64_closure;        // face projection, closure, eigenvalues, HLL in double
};

// Workgroup tile geometry of the FP32 stage kernels (cells per tile in x, y,
// z) of this translation unit; the batched kernels are compiled once per
// shape of M1_BATCH_SHAPES.
#ifndef M1_TILE_TX
#define M1_TILE_TX 16
#endif
#ifndef M1_TILE_TY
#define M1_TILE_TY 16
#endif
constexpr int F32_TX = M1_TILE_TX, F32_TY = M1_TILE_TY, F32_TZ = 256 / M1_TILE_TY;
// Tile-skip dilation: a step's three stages reach 3 x 3 = 9 cells (the WENO-Z
// stencil half-width per stage), so tiles within 9 cells of a tile holding
// phi above the threshold are evaluated.
constexpr int TILE_REACH = 9;
constexpr int tile_radius(int t) { return (TILE_REACH + t - 1) / t; }
// Face slabs of a tile, written by STAGE2 when asked: for axis a (x, y, z)
// and side s (0: low, 1: high), the largest phi within TILE_REACH cells of
// that face (slot 2a + s) and within TILE_GUARD cells (slot 6 + 2a + s).
constexpr int TILE_GUARD = 4;
constexpr int FACE_SLOTS = 12;

// Upper bound of phi over the cells of tile u lying within TILE_REACH
// (base 0) or TILE_GUARD (base 6) cells of tile t, where u sits at tile
// offset e = (ex, ey, ez) from t and f holds u's face slabs: the region is the
// intersection of one slab per nonzero axis (the low side of u when e > 0).
// Offsets of two or more tiles use the TILE_REACH slab (a superset of the
// region) and hold nothing within TILE_GUARD (every tile is >= 8 cells wide).
__device__ inline float facing_bound(const float* f, int ex, int ey, int ez, int base)
{
    const int e[3] = {ex, ey, ez};
    float r = 3.0e38f;
    for (int k = 0; k < 3; ++k) {
        if (e[k] == 0)
            continue;
        if (base != 0 && (e[k] > 1 || e[k] < -1))
            return 0.f;
        r = fminf(r, f[base + 2 * k + (e[k] > 0 ? 0 : 1)]);
    }
    return r;
}

struct TileMaskArgs {
    int ntx, nty, ntz;       // tiles per axis
    const float* tile_max;   // max phi per tile after this step (valid where computed)
    const float* tile_face;  // face slabs per tile (valid where computed)
    const uint8_t* computed; // tiles evaluated this step
    uint8_t* next;           // tiles to evaluate next step
    float threshold;         // activity threshold on the normalised phi
    int inflow_on;           // z = 0 tiles fed by the inflow
    int rx, ry, rz;          // dilation radius in tiles
    int* n_active;           // count of next-step tiles (atomic)
};

struct TileZeroArgs {
    GridDims g;
    const uint8_t* computed;
    const uint8_t* next;
    float* U;
    float* V1;
    float* V2;
};

struct CflArgs32 {
    GridDims g;
    const float* U;
    const uint8_t* ids;
    const double* spatial;
    const StepCoef* coef;
    unsigned long long* out;
    int wave_mode;
};

struct BoundaryArgs32 {
    GridDims g;
    const double* spatial;
    const uint8_t* ids;
    const StepCoef* coef;
    float* bnd;
    int integrating_factor;
};

// ------------------------------------------------ batched beamlets (FP32 bulk)
// Every beamlet marches in its own box of the phantom.  The boxes' arrays are
// concatenated: box b owns cells [cell_off, cell_off + n_b) of every per-cell
// array (the state holds 4 component planes of n_b cells from 4 * cell_off),
// columns [col_off, col_off + nx_b ny_b) of the inflow profile (boundary
// traces: 12 floats per column from 12 * col_off) and the workgroup tiles
// [tile_off, tile_off + ntx ntz nty) of the global tile index space.
struct BoxDesc {
    int nx, ny, nz;
    int ntx, nty, ntz;
    int tile_off;
    int pad;
    long cell_off;
    long col_off;
};

// Per box and step, written by the host.
struct BoxStep {
    int active;     // the beamlet is still marching
    int inflow_on;  // z = 0 tiles stay evaluated: the inflow spectrum is on
    int model_on;   // the tile-skip network may drop tiles of this box
    float energy;   // energy at the start of the next step, MeV
    int full_set;   // every tile of the box is evaluated next step
};

struct BatchStageArgs {
    const BoxDesc* box;
    const int* tile_box;     // global tile -> box
    const int* list;         // tiles to evaluate (global tile index), ascending
    int n_list;
    double dx, dy, dz;
    const float* src;
    float* U;
    const float* V1;
    float* out;
    float* dose_hi;
    float* dose_lo;
    const uint8_t* ids;
    const float* bnd;
    int bnd_stage;           // boundary traces of stage 0, 1 or 2
    const StepCoef32* coef;  // per box
    float* tile_max;         // per global tile (nullable)
    float* tile_sums;        // per global tile: sum phi, sum |J| (nullable)
    float* tile_face;        // per global tile: FACE_SLOTS slabs (nullable)
    int wave_mode;
    int f64_closure;         // face projection, closure, eigenvalues, HLL in double
    // stage 2 of boxes flagged in judge (nullable) adds its dose to
    // judge_hi / judge_lo instead of dose_hi / dose_lo
    const uint8_t* judge;
    float* judge_hi;
    float* judge_lo;
    // Persistent launch (queue not null): a fixed grid whose workgroups take
    // list entries from per-XCD queues (count[0] entries; the list split into
    // XCD_QUEUES contiguous slices, queue[x] counting the entries of slice x
    // handed out), stealing from the other slices once their own is empty.
    const int* count;
    int* queue;
    // Estimated launch (strided > 0, queue null): strided (a multiple of 8)
    // is an estimate of the list length count[0], which the kernel reads;
    // workgroup i < strided takes slot i (or exits past the list), and the
    // n_list - strided workgroups after them take the entries past the
    // estimate, grid-stride.
    int strided;
    // Stage 1 / 2 stencil input formed by the previous stage (win), and the
    // one this stage forms for the next (wout); nullable.
    const float* win;
    float* wout;
    // Relaxation sweeps (STAGE_RELAX): the quasi-steady cells are read from
    // V1 (the previous sweep's output; null: from src) and written to out,
    // the other cells read from src.  With convergence flags (relax_done_in
    // not null), a box is done for this sweep when it was done or the
    // previous sweep's largest change of its quasi-steady cells
    // (relax_change_prev, ordered float bits) is below relax_tol: its tiles
    // copy their quasi-steady cells from V1 to out; the others fold this
    // sweep's largest change into relax_change_cur.  Every workgroup of a box
    // also writes the box's flag for the next sweep (relax_done_out) and
    // clears its slot of the sweep after it (relax_change_clear).
    const uint8_t* relax_done_in;
    uint8_t* relax_done_out;
    const unsigned* relax_change_prev;
    unsigned* relax_change_cur;
    unsigned* relax_change_clear;
    float relax_tol;
};

// Per list: its length followed by one set of XCD queue counters per stage
// (the compaction that writes the list clears them).
constexpr int XCD_QUEUES = 8;
constexpr int LIST_HEADER = 1 + 3 * XCD_QUEUES;
// The compaction's per-chunk counts follow the header (chunks of 1024 flags).
constexpr int COMPACT_MAX_BLOCKS = 4096;
constexpr int LIST_HEADER_ALLOC = LIST_HEADER + COMPACT_MAX_BLOCKS;

struct BatchCflArgs {
    const BoxDesc* box;
    const BoxStep* step;
    const int* tile_box;
    int n_tiles;
    double dx, dy, dz;
    const float* U;
    const uint8_t* ids;
    const double* spatial;
    const uint8_t* live;       // per tile: state may be non-zero (else exact vacuum)
    const uint8_t* tile_mats;  // per tile: bit m set when material m occurs in it
    const StepCoef* coef;      // per box: S0, ed[0], rf, inv_alpha, eig_eps
    unsigned long long* out;   // per box, ordered bits of the maximum
    int wave_mode;
    int vacuum_aware;          // directional masks from active cells (EXACT speeds only)
    float act_threshold;       // active: normalised phi above this
    // Live-list form (plain CFL; null list: one workgroup per tile of every
    // box): the live tiles from the compacted list (count entries), and per
    // box the OR of tile_mats over its tiles outside the live set, whose
    // vacuum values the host folds in.
    const int* list;
    const int* count;
    unsigned long long* box_mats;
    int n_boxes;
    int max_columns;           // largest nx * ny of the boxes
    int tail_floor;            // cells with normalised phi <= act_threshold take the
                               // zero state's speeds (plain CFL)
};

struct BatchBoundaryArgs {
    const BoxDesc* box;
    const BoxStep* step;
    int n_boxes;
    const double* spatial;
    const uint8_t* ids;
    const StepCoef* coef;      // per box
    float* bnd;
    int integrating_factor;
};

struct BatchMaskArgs {
    const BoxDesc* box;
    const BoxStep* step;
    const int* tile_box;
    int n_tiles;
    const float* tile_max;
    const float* tile_face;
    const uint8_t* computed;
    uint8_t* next;
    float threshold;
    int rx, ry, rz;
    int all_on;               // no skipping: every tile of an active box
    uint8_t* pred;            // in/out (nullable): steps a tile stays left out
                              // (set by the network, counted down here)
    uint8_t* ring;            // out (nullable): tiles the network may leave out, 1 when
                              // no cell within TILE_GUARD cells is above the
                              // threshold, 2 when one is (guard 0 only)
    float* wmax;              // in/out (nullable): running max of tile_max
    const uint8_t* only;      // (nullable) boxes not flagged here get no tile
    int guard;                // 1: candidates keep TILE_GUARD cells from any cell above
                              // the threshold; 0: every tile the reach rule adds
};

// Tile features of the tile-skip network (11 per tile):
//   log10(max phi)/10, log10(mean phi)/10, sum|J|/sum phi, mean density/1000,
//   E/250, log10(max phi of the 6 face neighbours)/10 (-x +x -y +y -z +z);
// phi is the normalised state, log10 of 0 is taken as -30.
struct BatchFeatureArgs {
    const BoxDesc* box;
    const BoxStep* step;
    const int* tile_box;
    int n_tiles;
    const uint8_t* computed;
    const float* tile_max;
    const float* tile_sums;
    const float* tile_face;
    const float* tile_rho;    // mean density per tile, kg/m^3
    float* features;          // out (n_tiles, 11) (nullable)
    // inference (nullable model), on the tiles evaluated in the last step:
    // pred[t] = horizon when p_empty > decision, else 0; never set for a tile
    // above the threshold or a z = 0 tile fed by the inflow
    const uint16_t* w16;
    const float* params;
    float decision;
    float threshold;
    uint8_t* pred;
    int horizon;
    const int* gate;          // (nullable) the kernel does nothing while *gate == 0
};

// Recording: append (features, label) of the ring tiles saved `horizon`
// steps ago; label 1 when the tile stayed at or below the threshold.
struct BatchLabelArgs {
    int n_tiles;
    const uint8_t* cand;
    const float* features;
    const float* wmax;
    float threshold;
    float* out_x;
    float* out_y;
    int* count;
    int cap;
};

// Quasi-steady materials (m1_batch_set_quasi_steady): the tiles holding one
// of them, and the commit of a relaxation sweep.
struct BatchAirArgs {
    const BoxDesc* box;
    const BoxStep* step;
    const int* tile_box;
    int n_tiles;
    const uint8_t* live;       // tiles evaluated this step
    const uint8_t* tile_mats;  // per tile: bit m set when material m occurs in it
    uint32_t qs_mask;
    uint8_t* air;              // out: live tiles of active boxes holding a quasi-steady material
    // commit: the listed tiles' quasi-steady cells take `from` into `to`
    const int* list;
    const int* count;
    const uint8_t* ids;
    const float* from;
    float* to;
    // reset of the sweeps' convergence flags: done[b] = 0 (the flags the
    // first sweep reads), change[b] = 0 (the slot it fills) and
    // change[2 n_boxes + b] = +inf (the slot it reads as the change before it)
    int n_boxes;
    uint8_t* done;
    unsigned* change;
};

// Judge: per box, the largest |full - skip| dose increment and the largest
// full increment of one step (ordered float bits).
struct BatchCompareArgs {
    const BoxDesc* box;
    const uint8_t* judge;     // boxes to compare
    int n_boxes;
    const float* full_hi;
    const float* full_lo;
    const float* skip_hi;
    const float* skip_lo;
    unsigned* max_diff;
    unsigned* max_full;
};

// Per-box slice operations of the judge, on the boxes with sel[b] != 0:
//   SLICE_SAVE     state and dose -> checkpoint sel[b] - 1
//   SLICE_RESTORE  checkpoint sel[b] - 1 -> state and dose (sel[b] == 3: zero)
//   SLICE_RESET    judge doses (full and skip pairs) <- 0
//   SLICE_ADD      dose pair += skip pair
enum { SLICE_SAVE = 0, SLICE_RESTORE = 1, SLICE_RESET = 2, SLICE_ADD = 3 };
struct BatchSliceArgs {
    const BoxDesc* box;
    const uint8_t* sel;
    int n_boxes;
    int op;
    float* U;
    float* dose_hi;
    float* dose_lo;
    float* ck_U[2];
    float* ck_hi[2];
    float* ck_lo[2];
    float* full_hi;
    float* full_lo;
    float* skip_hi;
    float* skip_lo;
};

struct BatchZeroArgs {
    const BoxDesc* box;
    const int* tile_box;
    int n_tiles;
    const uint8_t* computed;
    const uint8_t* next;
    float* U;
    float* V1;
    float* V2;
    float* W1;  // the stage inputs formed by stages 0 and 1 (nullable)
    float* W2;
    float* R1;  // the quasi-steady relaxation's sweep outputs (nullable)
    float* R2;
};

// Launchers of the batched kernels compiled for one tile shape.
struct BatchKernels {
    int tx, ty, tz;  // cells per tile
    hipError_t (*stage)(int kind, bool integrating_factor, const BatchStageArgs& a,
                        hipStream_t s);
    hipError_t (*cfl)(const BatchCflArgs& a, hipStream_t s);
    hipError_t (*boundary)(const BatchBoundaryArgs& a, int max_columns, hipStream_t s);
    hipError_t (*mask)(const BatchMaskArgs& a, hipStream_t s);
    // list = ascending indices i < n with flags[i] != 0; count[0] = their
    // number, count[1 ..] (the stage queues) = 0
    hipError_t (*compact)(const uint8_t* flags, int n, int* list, int* count, hipStream_t s);
    hipError_t (*zero)(const BatchZeroArgs& a, hipStream_t s);
    // phantom[x0 + x, y0 + y, z] += hi + lo over the box (px, py, pz phantom shape)
    hipError_t (*scatter_dose)(const float* hi, const float* lo, int nx, int ny, int nz, int x0,
                               int y0, int py, int pz, double* phantom, hipStream_t s);
    hipError_t (*box_dose)(const float* hi, const float* lo, long n, double* out, hipStream_t s);
    hipError_t (*features)(const BatchFeatureArgs& a, hipStream_t s);
    hipError_t (*labels)(const BatchLabelArgs& a, hipStream_t s);
    hipError_t (*compare)(const BatchCompareArgs& a, int max_cells, hipStream_t s);
    hipError_t (*slices)(const BatchSliceArgs& a, int max_cells, hipStream_t s);
    // p[0 .. n) = 0 (a kernel, so that graphs hold no memset node)
    hipError_t (*clear)(unsigned long long* p, int n, hipStream_t s);
    // stage-kernel workgroups resident per CU (occupancy)
    int (*blocks_per_cu)();
    // launch form of a stage kernel (0 one tile per workgroup, 1 overflow,
    // 2 queue), -1 when func is not one of this table's stage kernels
    int (*stage_form)(const void* func);
    // quasi-steady tiles: the flags of the tiles to relax, and the commit of a
    // relaxation sweep over the listed tiles (grid: workgroups per list entry
    // estimate, each taking entries grid-stride)
    hipError_t (*air_mask)(const BatchAirArgs& a, hipStream_t s);
    hipError_t (*air_commit)(const BatchAirArgs& a, int grid, hipStream_t s);
    // before a series of sweeps: the convergence flags' reset (see BatchAirArgs)
    hipError_t (*air_reset)(const BatchAirArgs& a, hipStream_t s);
};

// Tile shapes (x, y, z cells; y * z = 256 threads, x * y <= 256) the batched
// kernels are compiled for, the default first.  The build reads this list.
#define M1_BATCH_SHAPES(X) X(16, 16, 16) X(16, 8, 32) X(8, 16, 16) X(32, 8, 32)
#define M1_BATCH_TABLE_NAME(a, b, c) batch_kernels_##a##x##b##x##c
#define M1_DECLARE_BATCH_TABLE(a, b, c) const BatchKernels& M1_BATCH_TABLE_NAME(a, b, c)();
M1_BATCH_SHAPES(M1_DECLARE_BATCH_TABLE)

// ------------------------------------------------ moving active domain (FP64)
// A box of the grid: half-open global cell ranges [i0, i1) x [j0, j1) x [k0, k1).
struct CellBox {
    int i0, i1, j0, j1, k0, k1;
};

// Reductions over the box state: the largest phi (order key, 0 = none) and
// the index range of the significant cells (lo = INT_MAX, hi = -1 = none).
struct BoxReduce {
    unsigned long long phi_max;
    int lo[3];
    int hi[3];
};

// Order-preserving 64-bit key of a double (negative values below positive,
// -0 below +0, NaN of either sign at the ends); 0 is below every key.
__host__ __device__ inline unsigned long long order_key(double v)
{
    unsigned long long b;
    __builtin_memcpy(&b, &v, sizeof b);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double key_value(unsigned long long k)
{
    const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double v;
    __builtin_memcpy(&v, &b, sizeof v);
    return v;
}

namespace box {
// dst (4 components on box db) = src (on box sb) where the boxes overlap,
// zero elsewhere; crop, rebox and embed are this copy between boxes.
hipError_t launch_copy(const double* src, CellBox sb, double* dst, CellBox db, hipStream_t s);
// ids and inflow profile of box b, cut from the full grid (ny, nz: full grid);
// the profile is zero unless the box touches z = 0.
hipError_t launch_gather(const uint8_t* ids, const double* spatial, int ny, int nz, CellBox b,
                         uint8_t* box_ids, double* box_spatial, hipStream_t s);
// dose (full grid) += local (box b), local = 0.
hipError_t launch_scatter_dose(double* dose, int ny, int nz, CellBox b, double* local,
                               hipStream_t s);
// r = {phi_max: 0, lo: INT_MAX, hi: -1}
hipError_t launch_reset(BoxReduce* r, hipStream_t s);
// r->phi_max = max(r->phi_max, key(phi)) over n cells
hipError_t launch_phi_max(const double* phi, long n, BoxReduce* r, hipStream_t s);
// Significant cells of a box of dims g: phi > tol * scale, and at z = 0
// (inflow_on) spatial * ed > tol * scale, scale = max(phi_max, phi_in_max if
// inflow_on); their index range goes to r->lo / r->hi.
hipError_t launch_significant(const double* phi, GridDims g, const double* spatial, double ed,
                              int inflow_on, double phi_in_max, double tol, BoxReduce* r,
                              hipStream_t s);
// act = phi > tol * scale or (z = 0 and spatial * ed > tol * scale), scale =
// max(phi_max, phi_in_max); every cell active when scale == 0.
hipError_t launch_activity(const double* phi, GridDims g, const double* spatial, double ed,
                           double phi_in_max, double tol, const BoxReduce* r, uint8_t* act,
                           hipStream_t s);
}  // namespace box

namespace f32 {
hipError_t launch_stage(int kind, bool integrating_factor, const StageArgs32& a, hipStream_t s);
hipError_t launch_cfl(const CflArgs32& a, hipStream_t s);
hipError_t launch_boundary(const BoundaryArgs32& a, hipStream_t s);
// dose64 = double(hi) + double(lo)
hipError_t launch_finalize_dose(const float* hi, const float* lo, double* dose, long n,
                                hipStream_t s);
// state64 = alpha * double(U32)
hipError_t launch_widen_state(const float* U, double alpha, double* out, long n, hipStream_t s);
hipError_t launch_narrow_state(const double* U, double inv_alpha, float* out, long n,
                               hipStream_t s);
hipError_t launch_scale(long n, float* U, const uint8_t* ids, const StepCoef* coef,
                        int nuclear, int scattering, hipStream_t s);
hipError_t launch_tile_mask(const TileMaskArgs& a, hipStream_t s);
hipError_t launch_tile_zero(const TileZeroArgs& a, hipStream_t s);
}  // namespace f32

#define M1_DECLARE_LAUNCHERS(NS)                                                          \
    namespace NS {                                                                        \
    hipError_t launch_stage(int kind, bool integrating_factor, const StageArgs& a,        \
                            hipStream_t s);                                               \
    hipError_t launch_cfl(const CflArgs& a, hipStream_t s);                               \
    hipError_t launch_boundary(const BoundaryArgs& a, hipStream_t s);                     \
    hipError_t launch_scale(const ScaleArgs& a, hipStream_t s);                           \
    }

M1_DECLARE_LAUNCHERS(exact)
M1_DECLARE_LAUNCHERS(fast)

}  // namespace m1
