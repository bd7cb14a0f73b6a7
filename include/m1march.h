/*
 * m1march.h -- C ABI of the MI355X M1 energy march (libm1march.so).
 *
 * The library replaces the hot path of anis-el-hachmioui/proton-dose-m1:
 * src/proton_solver/numerics/energy.py:1497-1644 (energy_march) and its
 * callees (spatial.py:104-252, boundary.py:117-399, weno_z.py:13-321,
 * flux.py:10-386, wave_speeds.py:12-105, realizability.py:132-215), the
 * moving active-domain march (energy.py:1262-1494 with
 * moving_active_domain.py:35-253) and the vacuum-aware CFL
 * (energy.py:459-537).
 * Signatures use plain C types only: the Python drop-in binds them with
 * ctypes (proton_solver/_native.py); any other FFI can bind them the same way.
 *
 * Arrays are C-ordered with z fastest, matching the reference's NumPy layout:
 *   state U      : (4, nx, ny, nz) float64, components (phi, Jx, Jy, Jz)
 *   dose         : (nx, ny, nz) float64 in Gy
 *   material ids : (nx, ny, nz) int32, values in [0, n_materials)
 *   inflow       : (nx, ny) float64 transverse fluence profile
 *                  (boundary.py:68-80, BeamInflow.spatial)
 *   boundary     : (4, nx, ny) float64 z-min inflow state
 *
 * Every function returns M1_OK (0) or a negative status; m1_last_error()
 * gives the message.  The Python layer raises ValueError for M1_ERR_VALUE and
 * NotImplementedError for M1_ERR_NOT_IMPLEMENTED, the exception types the
 * reference raises in the same situations.  A context owns one HIP stream and
 * its device buffers; it must be driven by one host thread at a time.
 */
#ifndef M1MARCH_H
#define M1MARCH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define M1_ABI_VERSION 5

enum {
    M1_OK = 0,
    M1_ERR_VALUE = -1,            /* invalid argument (ValueError)            */
    M1_ERR_NOT_IMPLEMENTED = -2,  /* unsupported option (NotImplementedError) */
    M1_ERR_HIP = -3,              /* HIP runtime error                        */
    M1_ERR_NONFINITE = -4,        /* march produced a non-finite value        */
    M1_ERR_STATE = -5             /* call order (no problem set, ...)         */
};

/* enums.py:22-38 (EnergyScheme subset the reference migrated, energy.py:1132-1165) */
enum { M1_FORWARD_EULER = 0, M1_SSPRK3 = 1, M1_IF_SSPRK3 = 2 };
/* enums.py:41-43 */
enum { M1_WAVE_EXACT = 0, M1_WAVE_UNIT = 1 };
/* enums.py:47-51 */
enum { M1_SCATTERING_OFF = 0, M1_SCATTERING_ROSSI = 1, M1_SCATTERING_WENTZEL = 2 };
/* enums.py:43-47 */
enum { M1_INTERP_LINEAR = 0, M1_INTERP_PCHIP = 1 };
/* Arithmetic of the march; no reference counterpart.  M1_FP64 keeps the
 * reference's operation order with every operation separately rounded (the
 * FP64 parity mode); M1_FP64_FAST lets the compiler fuse multiply-adds and
 * multiplies by reciprocal constants; M1_FP32_BULK is the mixed-precision
 * march (float state, reconstruction, closure, eigenvalues and fluxes, with
 * 1 - |J|^2/phi^2 formed exactly as |J|/phi -> 1, inflow extrapolation and
 * compensated dose in double / float pairs); M1_FP32_BULK_F64_CLOSURE is the
 * same with face projection, closure, eigenvalues and HLL in double. */
enum { M1_FP64 = 0, M1_FP32_BULK = 1, M1_FP64_FAST = 2, M1_FP32_BULK_F64_CLOSURE = 3 };
/* m1_step kinds: energy.py:663 forward_euler, 725 ssprk3, 878 if_ssprk3,
 * 1194 strang_energy_step. */
enum { M1_STEP_FORWARD_EULER = 0, M1_STEP_SSPRK3 = 1, M1_STEP_IF_SSPRK3 = 2,
       M1_STEP_STRANG = 3 };

typedef struct m1_ctx m1_ctx;

/* One tabulated coefficient (data/<table>/<material>.csv): n strictly
 * increasing energies [MeV], values, and PCHIP nodal slopes (required for the
 * stopping power, stopping_power.py:57-138; used for the transport and removal
 * tables only under M1_INTERP_PCHIP, may be NULL otherwise). */
typedef struct {
    int32_t n;
    const double *energy;
    const double *value;
    const double *slope;
} m1_table;

/* One material: physics/stopping_power.py, scattering.py, nuclear.py tables
 * and materials.py density / radiation length. */
typedef struct {
    m1_table stopping_power;   /* MeV/m */
    m1_table transport;        /* Wentzel first transport coefficient, 1/m */
    m1_table removal;          /* nuclear removal coefficient, 1/m */
    double density;            /* kg/m^3 */
    double radiation_length;   /* m */
} m1_material;

/* beam.py:6-13 fields the march reads (energy window and inflow spectrum). */
typedef struct {
    double energy;          /* MeV */
    double energy_spread;   /* MeV, must be > 0 (boundary.py:85-88) */
    double angular_spread;  /* rad */
} m1_beam;

/* config.py:16-41 fields the march reads, plus the precision knob. */
typedef struct {
    int32_t energy_scheme;              /* M1_FORWARD_EULER / M1_SSPRK3 / M1_IF_SSPRK3 */
    int32_t wave_speeds;                /* M1_WAVE_* */
    int32_t scattering;                 /* M1_SCATTERING_* */
    int32_t nuclear;                    /* 0 / 1 */
    int32_t coefficient_interpolation;  /* M1_INTERP_* */
    int32_t precision;                  /* M1_FP64 / M1_FP32_BULK / ... */
    double max_energy_step;             /* MeV */
    double cfl;
    double energy_min;                  /* MeV */
    double energy_window_sigma;
    /* FP32_BULK modes only: workgroup tiles (16 x 16 x 16 cells) whose 9-cell
     * neighbourhood holds phi below tile_threshold x (peak inflow fluence)
     * are not evaluated and are kept as exact vacuum; 0 evaluates every
     * tile.  No reference counterpart (the reference's closest analogue is
     * its moving active domain, moving_active_domain.py:107-193). */
    double tile_threshold;
    /* m1_march_moving: the vacuum-aware CFL of the moving active domain
     * (energy.py:459-537, 610-636, config.py:49) at the truncation tolerance.
     * Batched march with tile_threshold > 0: the same rule on the beamlet
     * boxes (ignored by m1_march, as the reference's static energy_march
     * ignores it).  A cell's axis speed counts only within 3 cells along
     * that axis of an active cell, active meaning normalised phi (or the z = 0
     * inflow phi) above tile_threshold; the reference's relative tolerance
     * is taken as the tile threshold, relative to the beamlet's peak inflow
     * fluence instead of the current maximum. */
    int32_t vacuum_aware_cfl;
    /* Batched march only: 0 bounds a beamlet's CFL step, as the march of
     * that beamlet on the whole phantom does, by the phantom outside its box
     * too (held at the zero state there; not with vacuum_aware_cfl, which
     * drops vacuum far from active cells); 1 bounds the box alone. */
    int32_t batch_cfl_box_only;
} m1_numerics;

/* Library ABI version (M1_ABI_VERSION). */
int m1_abi_version(void);

/* Create a context on HIP device `device`.  No reference counterpart: the
 * reference resolves its array module once per Solver (backend.py:12-59). */
int m1_create(m1_ctx **out, int device);
void m1_destroy(m1_ctx *ctx);

/* Message of the last failing call on ctx (or of the last m1_create failure
 * when ctx is NULL). */
const char *m1_last_error(const m1_ctx *ctx);

/* Bind a problem: grid (grid.py:19-42), material map (materials.py:65-127),
 * material tables, inflow profile (boundary.py:68-80) and beam.  Uploads
 * everything to device memory and zeroes the state and the dose.
 * Replaces the argument marshalling of energy_march (energy.py:1497-1505). */
int m1_set_problem(m1_ctx *ctx, int32_t nx, int32_t ny, int32_t nz,
                   double dx, double dy, double dz,
                   const int32_t *ids, int32_t n_materials,
                   const m1_material *materials,
                   const double *inflow_spatial, const m1_beam *beam);

/* Device state U <- host (4,nx,ny,nz) array; NULL sets U = 0
 * (solver.py:94-100). */
int m1_set_state(m1_ctx *ctx, const double *U);
int m1_get_state(m1_ctx *ctx, double *U);
/* Dose accumulated by the last m1_march / m1_step calls, Gy. */
int m1_get_dose(m1_ctx *ctx, double *dose);
int m1_zero_dose(m1_ctx *ctx);

/* energy_march (energy.py:1497-1644), static full domain: march the device
 * state from E0 = E_beam + energy_window_sigma * sigma_E down to energy_min.
 * The dose is reset, accumulated on the device and read with m1_get_dose; the
 * final state stays on the device (m1_get_state).  energies receives E0 and
 * the energy after every step (at most energies_cap values written; the full
 * count goes to *n_energies).  Snapshots (energy.py:1533-1541, 1567-1578,
 * 1626-1637): snapshots is n_snapshots*(4*nx*ny*nz) doubles, slot i filled
 * when snapshot_energies[i] was reached (snapshot_taken[i] = 1). */
int m1_march(m1_ctx *ctx, const m1_numerics *num,
             double *energies, int32_t energies_cap, int32_t *n_energies,
             const double *snapshot_energies, int32_t n_snapshots,
             double *snapshots, int32_t *snapshot_taken);

/* _moving_active_energy_march (energy.py:1262-1494, moving_active_domain=
 * True) with the box helpers of moving_active_domain.py:35-253.  The state
 * set with m1_set_state (zero for the reference's U = None) is cut to the
 * initial box (beam support widened by guard_cells, z in [0, 1 + guard_cells));
 * before every step the box becomes the bounding box of the cells with
 * phi > exp(-truncation_sigma^2 / 2) * max phi (the inflow phi included while
 * the box touches z = 0) widened by guard_cells, the state is moved to it
 * (zero where it is new), and the step runs on the box alone with the box's
 * materials and an inflow that is zero once the box has left z = 0.  With
 * num->vacuum_aware_cfl the energy step uses the vacuum-aware CFL at the
 * same tolerance (m1_compute_energy_step_vacuum).  The dose is accumulated on
 * the full grid (m1_get_dose); the final state, zero outside the last box,
 * is left on the device (m1_get_state; store_final_state=False is the
 * caller's choice not to read it).  energies, snapshots (full-grid states)
 * and their arguments as m1_march.  boxes (nullable) receives for each of
 * the first energies_cap steps the box {i0, i1, j0, j1, k0, k1} (half-open
 * global cell ranges) the step was taken on.  FP64 and FP64_FAST only;
 * guard_cells >= 4 and truncation_sigma > 0 (config.py:81-88). */
int m1_march_moving(m1_ctx *ctx, const m1_numerics *num, int32_t guard_cells,
                    double truncation_sigma, double *energies, int32_t energies_cap,
                    int32_t *n_energies, const double *snapshot_energies,
                    int32_t n_snapshots, double *snapshots, int32_t *snapshot_taken,
                    int32_t *boxes);

/* compute_energy_step with vacuum_relative_tolerance = relative_tolerance
 * (energy.py:493-660) of the current device state on the whole grid: an axis's
 * exact M1 speed counts in a cell only when a cell within 3 cells along that
 * axis has phi > relative_tolerance * max(max phi, max inflow phi), the inflow
 * phi counting at z = 0 (every cell counts when that maximum is zero).  With
 * unit wave speeds it equals m1_compute_energy_step.  FP64 and FP64_FAST. */
int m1_compute_energy_step_vacuum(m1_ctx *ctx, const m1_numerics *num, double energy,
                                  double relative_tolerance, double *dE);

/* Gamma index on the device: no reference counterpart (the reference's
 * metrics.py:542-606 calls pymedphys).  Counts the points (x, y, z voxel
 * triplets, n_points of them) whose global gamma^2 is <= 1: for each point p,
 * over the search offsets (offsets: n_offsets voxel-unit triplets, sorted so
 * that dist_terms = (length / distance)^2 is non-decreasing), the evaluated
 * dose is trilinearly interpolated at p + offset and gamma^2 = min
 * ((D - D_ref(p)) / dose_tolerance)^2 + dist_term over the offsets whose
 * p + offset lies inside the grid (pymedphys evaluates the dose outside the
 * grid as +inf, so such search points never match).  dose and
 * reference are host (nx, ny, nz) arrays; device memory is allocated and
 * freed inside the call.  The arithmetic is that of
 * proton_solver.metrics.gamma_pass_rate, so the count equals the host
 * search's.  m1_gamma_last_error gives the message of the calling thread's
 * last failed call. */
int m1_gamma_pass_count(int32_t device, const double *dose, const double *reference,
                        int32_t nx, int32_t ny, int32_t nz, const int32_t *points,
                        int32_t n_points, const double *offsets, const double *dist_terms,
                        int32_t n_offsets, double dose_tolerance, int64_t *n_pass);
const char *m1_gamma_last_error(void);

/* No counterpart in the reference: m1_gamma_pass_count for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  dose, reference (nx, ny, nz)
 * and points (n_points x 3) are DEVICE pointers the caller owns; offsets
 * and dist_terms are host arrays as in m1_gamma_pass_count and are staged,
 * with the pass counter, in a fixed device array of the library (about
 * 512 KiB per GPU, held until the process ends), so the call allocates
 * nothing.  n_offsets above M1_GAMMA_NOALLOC_MAX_OFFSETS is rejected
 * (M1_ERR_VALUE); points outside the grid are skipped by the kernel and
 * reported as M1_ERR_VALUE.  The work is queued on `stream` (a hipStream_t;
 * NULL: the default stream), which the call synchronizes before returning
 * the count.  Every call on one GPU shares that array: calls that overlap on
 * one GPU (several host threads, or streams whose calls are in flight at
 * once) race and must be serialized by the caller. */
#define M1_GAMMA_NOALLOC_MAX_OFFSETS 16384
int m1_gamma_pass_count_noalloc(int32_t device, const double *dose, const double *reference,
                                int32_t nx, int32_t ny, int32_t nz, const int32_t *points,
                                int32_t n_points, const double *offsets,
                                const double *dist_terms, int32_t n_offsets,
                                double dose_tolerance, void *stream, int64_t *n_pass);

/* spatial_operator (spatial.py:104-252, active_domain=False) of the current
 * device state at `energy`: rhs = -div F_HLL, (4,nx,ny,nz).  boundary_state
 * is the z-min inflow state (4,nx,ny) or NULL for inflow.state(energy). */
int m1_spatial_operator(m1_ctx *ctx, const m1_numerics *num, double energy,
                        const double *boundary_state, double *rhs);

/* compute_energy_step (energy.py:539-660, vacuum_relative_tolerance=None) of
 * the current device state. */
int m1_compute_energy_step(m1_ctx *ctx, const m1_numerics *num,
                           double energy, double *dE);

/* One integrator step of the current device state (kind = M1_STEP_*):
 * forward_euler (energy.py:663-722), ssprk3 (725-873), if_ssprk3 (878-1130)
 * or strang_energy_step (1194-1260).  The energy deposition of the step
 * (MeV per m^3 per unit fluence, the `energy_deposition` array) is written to
 * edep when it is not NULL; the device dose is not touched. */
int m1_step(m1_ctx *ctx, const m1_numerics *num, int32_t kind,
            double energy, double dE, double *edep);

/* Kernel timing with HIP events on the context's stream (no reference
 * counterpart; the reference times the whole march, solver.py:102-114). */
int m1_set_profiling(m1_ctx *ctx, int32_t enable);
/* Form of the FP64 / FP64_FAST stage kernels (no reference counterpart):
 * split = 1 (the default) runs each stage as a stencil-input pass, one face
 * flux pass per axis and an update pass, each at several waves per SIMD;
 * split = 0 runs the fused tile kernel (all of a stage in one launch, one
 * wave per SIMD).  Both give the same bits; the split form holds 16 more
 * doubles per cell of device memory. */
int m1_set_stage_form(m1_ctx *ctx, int32_t split);
int32_t m1_kernel_count(void);
const char *m1_kernel_name(int32_t kernel);
int m1_kernel_stats(m1_ctx *ctx, int32_t kernel, int64_t *launches,
                    double *total_ms);
int m1_reset_stats(m1_ctx *ctx);
/* Tiles evaluated by the stage kernels since the last m1_reset_stats, and the
 * number a full-domain march would have evaluated. */
int m1_tile_stats(m1_ctx *ctx, int64_t *evaluated, int64_t *full);
/* Cells per FP32-bulk workgroup tile along x, y and z of the single-box
 * kernels (fixed when the library is compiled). */
void m1_tile_shape(int32_t *tx, int32_t *ty, int32_t *tz);

/* ------------------------------------------------------------------
 * Batched beamlets.  No reference counterpart: the reference solves one
 * beam per energy_march call (solver.py:102-114).  A batch holds a
 * phantom (material ids on a regular grid) and many pencil beamlets that
 * enter at the phantom's z-min face; beamlet b is marched on its own box
 * of the phantom, cells [x0, x0+nx) x [y0, y0+ny) x [0, nz), with the
 * method of energy_march applied to that box as its domain (its own
 * energy grid from the CFL condition over the box, zero-gradient
 * boundaries on the box faces, inflow on its z-min face).  All boxes are
 * resident in device memory and each stage launch advances the evaluated
 * tiles of every beamlet still marching.  The batch runs the FP32-bulk
 * kernels (precision must be M1_FP32_BULK) with the IF-SSPRK3 scheme, or
 * SSPRK3 without scattering and nuclear removal.
 * ------------------------------------------------------------------ */
typedef struct m1_batch m1_batch;

typedef struct {
    int32_t x0, y0;        /* box origin in phantom cells */
    int32_t nx, ny, nz;    /* box extent in cells; the box starts at z = 0 */
    int32_t reserved;      /* 0 */
    m1_beam beam;
} m1_beamlet;

int m1_batch_create(m1_batch **out, int device);
void m1_batch_destroy(m1_batch *batch);
const char *m1_batch_last_error(const m1_batch *batch);

/* Phantom (nx, ny, nz) with cell sizes (dx, dy, dz) and ids in
 * [0, n_materials); beamlets[b] with its transverse inflow profile at
 * inflow_spatial + sum_{c<b} nx_c*ny_c, shape (nx_b, ny_b) (the
 * BeamInflow.spatial of the beam on the box grid, boundary.py:68-80). */
int m1_batch_set_problem(m1_batch *batch, int32_t nx, int32_t ny, int32_t nz,
                         double dx, double dy, double dz, const int32_t *ids,
                         int32_t n_materials, const m1_material *materials,
                         int32_t n_beamlets, const m1_beamlet *beamlets,
                         const double *inflow_spatial);

/* March every beamlet from E_b + energy_window_sigma * sigma_E down to
 * energy_min.  dose (phantom shape, Gy; may be NULL) receives the sum of
 * the beamlet doses; n_steps (n_beamlets; may be NULL) the number of energy
 * steps of each beamlet. */
int m1_batch_march(m1_batch *batch, const m1_numerics *num, double *dose,
                   int32_t *n_steps);

/* Dose of beamlet b from the last march on its box, (nx_b, ny_b, nz_b), Gy. */
int m1_batch_beamlet_dose(m1_batch *batch, int32_t beamlet, double *dose);

/* Kernel timing and tile counts, as m1_set_profiling / m1_kernel_stats /
 * m1_tile_stats / m1_reset_stats. */
int m1_batch_set_profiling(m1_batch *batch, int32_t enable);
int m1_batch_kernel_stats(m1_batch *batch, int32_t kernel, int64_t *launches,
                          double *total_ms);
int m1_batch_tile_stats(m1_batch *batch, int64_t *evaluated, int64_t *full);
int m1_batch_reset_stats(m1_batch *batch);
/* Host time of the per-step physics scalars since the last reset, ms. */
int m1_batch_host_ms(m1_batch *batch, double *setup_ms);

/* CFL reduction of the batch (no reference counterpart).  By default the
 * plain CFL evaluates the cells of the live tiles (those that may hold a
 * non-zero state) from the compacted tile list and takes the tiles at the
 * zero state by their materials (their isotropic speeds) and z-min inflow
 * columns; enable = 1 evaluates every cell of every tile of the boxes
 * instead.  Both give the same maximum bit for bit. */
int m1_batch_set_cfl_all_tiles(m1_batch *batch, int32_t enable);
/* Plain batch CFL with a tail floor (enable 1; no reference counterpart):
 * a cell whose phi is at or below tile_threshold (normalised, the level of
 * the tile set) counts with the speeds of the zero state, which is what it
 * holds when its tile is left out.  The energy steps then depend on the
 * cells above the threshold, the inflow and the materials present, not on
 * which sub-threshold tiles are evaluated; sub-threshold cells may step up
 * to ~1.2x past their own CFL bound (the largest M1 speed sum over the zero
 * state's). */
int m1_batch_set_cfl_tail_floor(m1_batch *batch, int32_t enable);

/* Stage kernels on a persistent grid (the workgroups resident on the
 * device) that take tiles from eight per-XCD queues, stealing across XCDs
 * when their own is empty (enable 1), or one workgroup per tile (0, the
 * default).  Results are identical; only the work distribution changes. */
int m1_batch_set_persistent(m1_batch *batch, int32_t enable);
/* Tile shapes the batched kernels are compiled for (cells per tile along x,
 * y, z; index 0 is the default, 16 x 16 x 16).  m1_batch_set_tile_shape
 * selects one for the batch and must precede m1_batch_set_problem.  No
 * reference counterpart; proton_solver/autotune.py times them. */
int32_t m1_batch_tile_shape_count(void);
int m1_batch_tile_shape(int32_t index, int32_t *tx, int32_t *ty, int32_t *tz);
int m1_batch_set_tile_shape(m1_batch *batch, int32_t index);
/* Capture a step's launches (step coefficients, boundary traces, the three
 * persistent stage kernels, the next tile set and the next CFL reduction)
 * as a HIP graph per tile-set parity and replay it for every step that is
 * neither judged nor recorded (enable 1), or launch them one by one (0, the
 * default).  Results are identical.  With kernel timing on
 * (m1_batch_set_profiling) every step is launched one by one. */
int m1_batch_set_graphs(m1_batch *batch, int32_t enable);
/* Steps of the last march replayed from a graph; *capture_error (nullable)
 * is empty, or why a capture failed and the march fell back to launches. */
int m1_batch_graph_info(m1_batch *batch, int64_t *graph_steps, const char **capture_error);
/* Device-side energy stepping (enable 1; no reference counterpart; replaces
 * the per-step host work of energy_march, numerics/energy.py:1556-1624, for
 * the batch): each beamlet's CFL step (energy.py:640-660), its step scalars
 * (stopping_power.py:141-228, scattering.py:267-306, nuclear.py:202-233,
 * energy.py:273-310, boundary.py:82-114), its energy, the judge's decisions
 * and the box flags are computed by kernels, and the host enqueues steps
 * without waiting for them (it reads each step's status a few steps later).
 * Sums, products, quotients and square roots round as the host's; the
 * exponentials are the device library's, so results agree with host-side
 * stepping to the last bits of the exponentials, not always bit for bit.
 * With the tile-skip network, beamlets are judged on a common schedule
 * (every judge_every-th step of the batch) rather than per beamlet, and not
 * on their last step.  Recording marches keep host-side stepping. */
int m1_batch_set_device_steps(m1_batch *batch, int32_t enable);
/* Progress of device-stepped marches (no reference counterpart): with
 * every > 0 a line "m1 batch: energy step k, n beamlets marching" goes to
 * stderr every `every` energy steps (0, the default: none). */
int m1_batch_set_progress(m1_batch *batch, int32_t every);
/* Quasi-steady materials (no reference counterpart; for phantoms holding
 * air, whose stopping power is 1/934 of water's so that an explicit energy
 * step over air is 934 times shorter).  Materials whose bit is set in
 * material_mask hold their state through each stage of an energy step
 * (their flux divergence is not applied and they set no CFL bound); after
 * each stage `iterations` Jacobi sweeps of a pseudo-time step towards
 * div F = 0 (forward Euler, pseudo-time CFL number pseudo_cfl with each
 * cell's own M1 speeds) relax their cells of the next stage's input with the
 * other cells fixed: the beam crosses them at no energy cost, the limit of
 * the reference's march as their stopping power goes to 0.  Their dose is
 * deposited from the held and relaxed states.  Needs device-side stepping;
 * not combined with the tile-skip network.  iterations 0 turns it off (the
 * default). */
int m1_batch_set_quasi_steady(m1_batch *batch, uint32_t material_mask, int32_t iterations,
                              double pseudo_cfl);
/* Convergence of the quasi-steady sweeps (no reference counterpart): with
 * tolerance > 0 a beamlet stops sweeping after a stage once a sweep changed
 * none of its quasi-steady cells by more than tolerance times its peak
 * inflow fluence (at most `iterations` sweeps); 0, the default, runs every
 * sweep.  Tolerance in [0, 1). */
int m1_batch_set_quasi_steady_tolerance(m1_batch *batch, double tolerance);
/* Device evaluation of the step physics of the bound problem's material
 * `material` at n energies (test access; no reference counterpart): kind 0
 * S(E) (stopping_power.py:141-228), 1 the transport coefficient of
 * `scattering` (scattering.py:65-153), 2 the removal coefficient
 * (nuclear.py:94-138), 3 the scattering optical depth over [E - dE, E]
 * (scattering.py:267-306), 4 the nuclear optical depth (nuclear.py:202-233);
 * interpolation as m1_numerics.coefficient_interpolation. */
int m1_batch_physics_probe(m1_batch *batch, int32_t material, int32_t kind, int32_t scattering,
                           int32_t interpolation, const double *energies, double dE, int32_t n,
                           double *out);

/* ------------------------------------------------------------------
 * Tile-skip network (no reference counterpart): 11 inputs per workgroup
 * tile, three hidden layers of 64 ReLU units, one logit (probability that
 * the tile stays empty over the next steps).  Weights are applied in
 * bfloat16 (rounded to nearest even from the float32 master copy), biases
 * and accumulations in float32.  Parameters are a flat float32 array of
 * m1_mlp_param_count() values:
 *   W1[11][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4
 * (W[i][j]: input i to unit j).  Training (Adam on the mean binary
 * cross-entropy, shuffled minibatches) runs on the device.
 * ------------------------------------------------------------------ */
typedef struct m1_mlp m1_mlp;

#define M1_MLP_INPUTS 11

int32_t m1_mlp_param_count(void);
int m1_mlp_create(m1_mlp **out, int device);
void m1_mlp_destroy(m1_mlp *model);
const char *m1_mlp_last_error(const m1_mlp *model);
/* Set / read the float32 master parameters (setting resets Adam's moments). */
int m1_mlp_set_params(m1_mlp *model, const float *params);
int m1_mlp_get_params(m1_mlp *model, float *params);
/* Probabilities sigmoid(logit) of n samples x (n, 11). */
int m1_mlp_forward(m1_mlp *model, const float *x, int32_t n, float *prob);
/* Mean binary cross-entropy of labels y (0 or 1) and its gradient with
 * respect to the parameters (weights entering through their bf16 values). */
int m1_mlp_loss_grad(m1_mlp *model, const float *x, const float *y, int32_t n,
                     double *loss, float *grad);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over `epochs` passes of shuffled
 * minibatches of `batch` samples; *final_loss is the mean loss over the last
 * pass. */
int m1_mlp_train(m1_mlp *model, const float *x, const float *y, int32_t n,
                 int32_t epochs, int32_t batch, float lr, uint64_t seed,
                 double *final_loss);

/* ------------------------------------------------------------------
 * Tile-skip network inside the batched march (no reference counterpart).
 *
 * Features of a tile (computed on the device from the last step):
 *   0  log10(max phi) / 10          1  log10(mean phi) / 10
 *   2  sum |J| / sum phi            3  mean density / 1000 (g/cm^3)
 *   4  beamlet energy / 250 MeV     5-10  log10(max phi) / 10 of the six
 *                                         face neighbours (-x +x -y +y -z +z)
 * (phi below 1e-30 counts as 1e-30; tiles not evaluated count as empty).
 *
 * Label: 1 when a candidate tile (below) keeps phi at or below the
 * threshold over the next `horizon` steps.
 *
 * With a model set, the network is evaluated every `infer_every` steps
 * (1..horizon) on the tiles that step evaluated; a tile whose probability
 * exceeds `decision` is left out for the next `horizon` steps (1..255)
 * whenever it is a candidate: the reach rule adds it while no cell within
 * 4 cells of it holds phi above the threshold (checked every step).  Tiles
 * above the threshold and inflow tiles are never left out.
 * Every `judge_every` (0: never)
 * steps of a beamlet, and on its last step, the judge takes that step twice
 * from the same state: on every tile of the beamlet's box and on the tile set
 * in use.  When the two step doses differ anywhere by more than judge_tol
 * times the step's largest dose increment, the beamlet goes back to the state
 * and dose of its last accepted judge (or to its start) and marches on from
 * there without the network; otherwise that state becomes its checkpoint.
 * The model must stay alive, unchanged, while the batch uses it; marches
 * need tile_threshold > 0.
 * ------------------------------------------------------------------ */
/* model NULL switches the network off. */
int m1_batch_set_skip_model(m1_batch *batch, const m1_mlp *model, float decision,
                            int32_t horizon, int32_t infer_every, int32_t judge_every,
                            float judge_tol);
/* Candidates of the network: guard_cells = 4 (the default) admits the tiles
 * the reach rule adds while no cell within 4 cells of them holds phi above
 * the threshold; 0 admits every tile the reach rule adds that holds no such
 * cell itself (the network alone then decides on the tiles next to the
 * beam).  Applies to the marches and to the recording below. */
int m1_batch_set_skip_guard(m1_batch *batch, int32_t guard_cells);
/* Record (features, label) samples of the dilation-only tiles during the
 * following marches, every `horizon` steps, up to `capacity` samples.
 * full_domain: evaluate every tile of the active boxes (labels from the
 * unskipped march).  enable 0 stops recording and frees the samples. */
int m1_batch_record(m1_batch *batch, int32_t enable, int32_t full_domain,
                    int32_t horizon, int64_t capacity);
/* *count = samples held; copies min(count, capacity) of them into x
 * (count, 11) and y: the label (1: the tile stayed at or below the threshold
 * over the horizon), plus 2 when a cell within 4 cells of the tile was above
 * the threshold when the sample was taken (possible with guard 0 only). */
int m1_batch_samples(m1_batch *batch, float *x, float *y, int64_t capacity,
                     int64_t *count);
/* Judge checks made and beamlets rejected since the model was set;
 * rollback_to (nullable) receives, per beamlet, the step count it was sent
 * back to in the last march (-1: never rejected). */
int m1_batch_skip_stats(m1_batch *batch, int64_t *judged, int64_t *rejections,
                        int32_t *rollback_to);


#ifdef __cplusplus
}
#endif
#endif
