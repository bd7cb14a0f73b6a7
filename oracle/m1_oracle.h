/*
 * m1_oracle.h -- CPU FP64 restatement of the reference M1 energy march.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in proton-dose-m1_amd/csrc.  Only tests/, the smoke() entry
 * of __graft_entry__.py and the cpu_baseline leg of bench.py may load it, and
 * only as the checker or the timed CPU baseline -- never as the product.
 *
 * Every function restates one function of the reference Python/Numba code
 * (anis-el-hachmioui/proton-dose-m1, src/proton_solver/...) literally: the same
 * operation order, clamps and constants, unfused, one pass per reference
 * kernel.  Citations are given next to each declaration.
 */
#ifndef M1_ORACLE_H
#define M1_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A tabulated coefficient: n strictly increasing energies, values and (for
 * PCHIP) nodal slopes.  slope may be NULL for linear interpolation. */
typedef struct {
    int n;
    const double *E;
    const double *v;
    const double *slope;
} or_table;

typedef struct {
    or_table sp;          /* electronic stopping power [MeV/m], PCHIP      */
    or_table tr;          /* Wentzel first transport coefficient [1/m]     */
    or_table nu;          /* nuclear removal coefficient [1/m]             */
    double density;       /* kg/m^3                                        */
    double radiation_length; /* m (Rossi scattering model)                 */
} or_material;

typedef struct {
    int nx, ny, nz;
    double dx, dy, dz;
    const int32_t *ids;   /* nx*ny*nz material ids, C order (z fastest)    */
    int n_mat;
    const or_material *mats;
    const double *spatial;   /* nx*ny inflow fluence profile (BeamInflow.spatial) */
    double beam_energy;      /* MeV */
    double energy_spread;    /* MeV */
    double angular_spread;   /* rad */
} or_problem;

enum { OR_FORWARD_EULER = 0, OR_SSPRK3 = 1, OR_IF_SSPRK3 = 2 };
enum { OR_WAVE_EXACT = 0, OR_WAVE_UNIT = 1 };
enum { OR_SCAT_OFF = 0, OR_SCAT_ROSSI = 1, OR_SCAT_WENTZEL = 2 };
enum { OR_INTERP_LINEAR = 0, OR_INTERP_PCHIP = 1 };

typedef struct {
    int energy_scheme;
    int wave_speeds;
    int scattering;
    int nuclear;
    int coefficient_interpolation;
    double max_energy_step;
    double cfl;
    double energy_min;
    double energy_window_sigma;
} or_config;

const char *or_last_error(void);
void or_set_threads(int n);

/* --- physics scalars (physics/ modules) ------------------------------ */
/* stopping_power.py:30-54 and 57-138: PCHIP nodal slopes. */
int or_pchip_slopes(int n, const double *E, const double *v, double *slope_out);
/* stopping_power.py:141-228 (range check with np.isclose, clamp, Hermite). */
int or_pchip_eval(const or_table *t, double energy, double *out);
/* numpy.interp on a clamped energy (scattering.py:93-108, nuclear.py:115-130). */
double or_linear_eval(const or_table *t, double energy);
int or_coefficient(const or_table *t, double energy, int interp, double *out);
int or_stopping_power(const or_material *m, double energy, double *out);
/* scattering.py:267-306 / nuclear.py:202-233: 2-point Gauss-Legendre. */
int or_scattering_depth(const or_material *m, int model, int interp,
                        double energy, double dE, double *out);
int or_nuclear_depth(const or_material *m, int interp,
                     double energy, double dE, double *out);

/* --- pointwise / stencil kernels (numerics/ modules) ----------------- */
/* weno_z.py:13-87 */
double or_weno_z_left(double v0, double v1, double v2, double v3, double v4);
/* realizability.py:132-176 (in place, n cells of a (4,n) array). */
void or_project_realizable(double *U, long n, double eps);
/* wave_speeds.py:12-105 */
void or_eigenvalues(const double *U, long n, int axis, double *lmin, double *lmax);
/* flux.py:137-386, faces in a (4,n) layout. */
void or_hll_flux(const double *UL, const double *UR, long n, int axis,
                 int wave_mode, double *F);
/* boundary.py:117-399 with ghost width 3; U_in is (4,nx,ny). */
int or_fill_ghost_cells(const double *U, int nx, int ny, int nz,
                        const double *U_in, double *U_ext);
/* weno_z.py:136-321 (reconstruct + per-face projection). axis 0,1,2. */
void or_reconstruct(const double *U_ext, int nx, int ny, int nz, int axis,
                    double *UL, double *UR);
/* spatial.py:104-252 (active_domain=False).  boundary_state (4,nx,ny) may
 * be NULL, in which case inflow.state(energy) is used. */
int or_spatial_operator(const or_problem *p, int wave_mode, const double *U,
                        double energy, const double *boundary_state,
                        double *rhs);
/* energy.py:180-202 / 412-424: stage-1 and stage-2 boundary values from
 * four samples of the (transformed) trace at tau/dE = 0, 1/3, 2/3, 1. */
void or_rk3_trace_combination(long n, const double *B0, const double *B13,
                              const double *B23, const double *B1,
                              double *Bs1, double *Bs2);
/* boundary.py:82-114 */
int or_inflow_state(const or_problem *p, double energy, double *U_in);
/* energy.py:539-660 (vacuum_relative_tolerance=None). */
int or_compute_energy_step(const or_problem *p, const or_config *c,
                           const double *U, double energy, double *dE_out);
/* energy.py:539-660 with vacuum_relative_tolerance given (energy.py:459-537):
 * the directional speeds of cells more than 3 cells (the WENO-Z stencil
 * reach) along that axis from any cell with phi > tolerance * max phi (the
 * inflow phi included at z = 0) are dropped. */
int or_compute_energy_step_vacuum(const or_problem *p, const or_config *c,
                                  const double *U, double energy,
                                  double relative_tolerance, double *dE_out);
/* energy.py:663-722, 725-873, 878-1130, 1194-1260: one step each.  U is
 * replaced by the new state, edep receives the energy deposition [MeV/m^3
 * per proton-fluence unit] of the step. */
int or_forward_euler(const or_problem *p, const or_config *c, double *U,
                     double energy, double dE, double *edep);
int or_ssprk3(const or_problem *p, const or_config *c, double *U,
              double energy, double dE, double *edep);
int or_if_ssprk3(const or_problem *p, const or_config *c, double *U,
                 double energy, double dE, double *edep);
int or_strang_step(const or_problem *p, const or_config *c, double *U,
                   double energy, double dE, double *edep);

/* energy.py:1497-1644 (static full domain).  U (4N) in/out, dose (N) out in
 * Gy.  energies receives E0 and the energy after every step (capacity cap);
 * n_energies the count.  Snapshots: n_snap requested energies; snap_out is
 * n_snap*4N; snap_taken[i] is set to 1 when snapshot i was stored.
 * max_steps > 0 stops the march after that many steps (bounded CPU-baseline
 * samples); 0 marches to energy_min. */
int or_energy_march(const or_problem *p, const or_config *c, double *U,
                    double *dose, double *energies, int cap, int *n_energies,
                    const double *snap_E, int n_snap, double *snap_out,
                    int *snap_taken, int max_steps);

/* energy.py:1262-1494 (moving_active_domain=True) with the box helpers of
 * moving_active_domain.py:35-253.  The state lives on a box that follows the
 * cells with phi > exp(-sigma^2/2) * max phi, widened by `guard` cells; the
 * dose is accumulated on the full grid.  U0 (full grid, 4N) may be NULL (zero
 * start on the initial box); U_out (4N, may be NULL) receives the final box
 * state embedded in the full grid; snapshots are full-grid states.  boxes
 * (may be NULL) receives per step the box {i0, i1, j0, j1, k0, k1} the step
 * was taken on (capacity cap steps).  Other arguments as or_energy_march. */
int or_energy_march_moving(const or_problem *p, const or_config *c, int guard,
                           double truncation_sigma, int vacuum_aware,
                           const double *U0, double *U_out, double *dose,
                           double *energies, int cap, int *n_energies,
                           const double *snap_E, int n_snap, double *snap_out,
                           int *snap_taken, int max_steps, int32_t *boxes);

#ifdef __cplusplus
}
#endif
#endif
