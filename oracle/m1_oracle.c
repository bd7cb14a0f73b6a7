/*
 * m1_oracle.c -- CPU FP64 restatement of the reference M1 energy march.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and timed CPU baseline).  Nothing
 * in the product links or loads this file; see m1_oracle.h.
 *
 * The code is deliberately unfused and structured like the reference's
 * Numba/NumPy pipeline: ghost fill -> per-axis WENO-Z reconstruction ->
 * per-face projection -> per-axis HLL -> divergence, then NumPy-style stage
 * algebra on full arrays.  Floating-point expressions keep the reference's
 * operation order; the file is compiled with -ffp-contract=off so that no
 * multiply-add is fused.  Reference paths are relative to
 * src/proton_solver/ of anis-el-hachmioui/proton-dose-m1.
 */
#include "m1_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define GHOST 3
#define MEV_TO_JOULE 1.602176634e-13           /* numerics/energy.py:37       */
#define EIG_EPS 1.0e-10                         /* numerics/wave_speeds.py:10  */
#define WENO_EPS 1.0e-12                        /* weno_z.py:8                 */
#define REAL_EPS 1.0e-10                        /* weno_z.py:10, realizability */
#define PROTON_REST_ENERGY_MEV 938.27208816     /* physics/scattering.py:20    */
#define ROSSI_ES_MEV 15.0                       /* physics/scattering.py:21    */

static char g_err[512];

const char *or_last_error(void) { return g_err; }

static int fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return -1;
}

void or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0)
        omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* physics scalars                                                     */
/* ------------------------------------------------------------------ */

/* numpy.isclose(a, b) with rtol=1e-5, atol=1e-8. */
static int np_isclose(double a, double b)
{
    return fabs(a - b) <= (1e-08 + 1e-05 * fabs(b));
}

static double sgn(double x) { return (x > 0.0) - (x < 0.0); }

/* physics/stopping_power.py:30-54 */
static double pchip_endpoint_slope(double h0, double h1, double d0, double d1)
{
    double slope = ((2.0 * h0 + h1) * d0 - h0 * d1) / (h0 + h1);
    if (sgn(slope) != sgn(d0))
        return 0.0;
    if (sgn(d0) != sgn(d1) && fabs(slope) > 3.0 * fabs(d0))
        return 3.0 * d0;
    return slope;
}

/* physics/stopping_power.py:93-138 (coefficient_interpolation.py:19-48). */
int or_pchip_slopes(int n, const double *E, const double *v, double *s)
{
    if (n < 2)
        return fail("table must contain at least two points");
    double *h = malloc(sizeof(double) * (size_t)(n - 1));
    double *d = malloc(sizeof(double) * (size_t)(n - 1));
    if (!h || !d) {
        free(h);
        free(d);
        return fail("out of memory");
    }
    for (int i = 0; i < n - 1; ++i) {
        h[i] = E[i + 1] - E[i];
        d[i] = (v[i + 1] - v[i]) / h[i];
    }
    for (int i = 0; i < n; ++i)
        s[i] = 0.0;
    if (n == 2) {
        s[0] = s[1] = d[0];
    } else {
        for (int k = 1; k < n - 1; ++k) {
            if (d[k - 1] * d[k] > 0.0) {
                double w1 = 2.0 * h[k] + h[k - 1];
                double w2 = h[k] + 2.0 * h[k - 1];
                s[k] = (w1 + w2) / (w1 / d[k - 1] + w2 / d[k]);
            }
        }
        s[0] = pchip_endpoint_slope(h[0], h[1], d[0], d[1]);
        s[n - 1] = pchip_endpoint_slope(h[n - 2], h[n - 3], d[n - 2], d[n - 3]);
    }
    free(h);
    free(d);
    return 0;
}

/* index i with E[i] <= x < E[i+1] (numpy searchsorted side='right' - 1). */
static int upper_index(const double *E, int n, double x)
{
    int lo = 0, hi = n; /* first index with E[idx] > x */
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (E[mid] <= x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo - 1;
}

static int range_check(const or_table *t, double energy)
{
    double e0 = t->E[0], e1 = t->E[t->n - 1];
    if ((energy < e0 && !np_isclose(energy, e0)) ||
        (energy > e1 && !np_isclose(energy, e1)))
        return fail("Energy %.17g MeV outside table range [%.17g, %.17g] MeV.",
                    energy, e0, e1);
    return 0;
}

static double clip(double x, double lo, double hi)
{
    return x < lo ? lo : (x > hi ? hi : x);
}

/* Hermite evaluation: stopping_power.py:171-228, coefficient_interpolation.py:51-70.
 * energy already clamped to the table. */
static double hermite(const or_table *t, double energy)
{
    int n = t->n;
    if (energy >= t->E[n - 1])
        return t->v[n - 1];
    int i = upper_index(t->E, n, energy);
    if (i < 0)
        i = 0;
    if (i > n - 2)
        i = n - 2;
    double h = t->E[i + 1] - t->E[i];
    double tt = (energy - t->E[i]) / h;
    double t2 = tt * tt;
    double t3 = t2 * tt;
    double h00 = 2.0 * t3 - 3.0 * t2 + 1.0;
    double h10 = t3 - 2.0 * t2 + tt;
    double h01 = -2.0 * t3 + 3.0 * t2;
    double h11 = t3 - t2;
    return h00 * t->v[i] + h10 * h * t->slope[i] + h01 * t->v[i + 1] +
           h11 * h * t->slope[i + 1];
}

int or_pchip_eval(const or_table *t, double energy, double *out)
{
    if (range_check(t, energy))
        return -1;
    energy = clip(energy, t->E[0], t->E[t->n - 1]);
    *out = hermite(t, energy);
    return 0;
}

/* numpy.interp (numpy/_core/src/multiarray/compiled_base.c, arr_interp) for
 * one abscissa inside [E0, En]. */
double or_linear_eval(const or_table *t, double x)
{
    int n = t->n;
    const double *dx = t->E, *dy = t->v;
    if (x < dx[0])
        return dy[0];
    if (x > dx[n - 1])
        return dy[n - 1];
    int j = upper_index(dx, n, x);
    if (j == n - 1)
        return dy[j];
    if (dx[j] == x)
        return dy[j];
    double slope = (dy[j + 1] - dy[j]) / (dx[j + 1] - dx[j]);
    double r = slope * (x - dx[j]) + dy[j];
    if (isnan(r)) {
        r = slope * (x - dx[j + 1]) + dy[j + 1];
        if (isnan(r) && dy[j] == dy[j + 1])
            r = dy[j];
    }
    return r;
}

/* scattering.py:65-116 / nuclear.py:94-138 */
int or_coefficient(const or_table *t, double energy, int interp, double *out)
{
    if (range_check(t, energy))
        return -1;
    energy = clip(energy, t->E[0], t->E[t->n - 1]);
    if (interp == OR_INTERP_LINEAR)
        *out = or_linear_eval(t, energy);
    else
        *out = hermite(t, energy);
    return 0;
}

int or_stopping_power(const or_material *m, double energy, double *out)
{
    return or_pchip_eval(&m->sp, energy, out);
}

/* scattering.py:118-153 */
static double rossi_coefficient(const or_material *m, double energy)
{
    double e = energy > 1.0e-8 ? energy : 1.0e-8;
    double mm = PROTON_REST_ENERGY_MEV;
    double pv = e * (e + 2.0 * mm) / (e + mm);
    double r = ROSSI_ES_MEV / pv;
    return r * r / m->radiation_length;
}

/* scattering.py:155-181 and 227-265 (one material) */
static int scattering_rate(const or_material *m, int model, int interp,
                           double energy, double *out)
{
    double sig, S;
    if (model == OR_SCAT_WENTZEL) {
        if (or_coefficient(&m->tr, energy, interp, &sig))
            return -1;
    } else if (model == OR_SCAT_ROSSI) {
        sig = rossi_coefficient(m, energy);
    } else {
        sig = 0.0;
    }
    if (or_stopping_power(m, energy, &S))
        return -1;
    *out = sig / S;
    return 0;
}

/* scattering.py:267-306 */
int or_scattering_depth(const or_material *m, int model, int interp,
                        double energy, double dE, double *out)
{
    if (model == OR_SCAT_OFF) {
        *out = 0.0;
        return 0;
    }
    double mid = energy - 0.5 * dE;
    double hw = 0.5 * dE;
    double off = hw / sqrt(3.0);
    double a, b;
    if (scattering_rate(m, model, interp, mid - off, &a) ||
        scattering_rate(m, model, interp, mid + off, &b))
        return -1;
    *out = hw * (a + b);
    return 0;
}

/* nuclear.py:163-233 */
int or_nuclear_depth(const or_material *m, int interp, double energy,
                     double dE, double *out)
{
    if (dE <= 0.0)
        return fail("dE must be positive.");
    double mid = energy - 0.5 * dE;
    double hw = 0.5 * dE;
    double off = hw / sqrt(3.0);
    double ea = mid - off, eb = mid + off, sa, sb, Sa, Sb;
    if (or_coefficient(&m->nu, ea, interp, &sa) || or_stopping_power(m, ea, &Sa) ||
        or_coefficient(&m->nu, eb, interp, &sb) || or_stopping_power(m, eb, &Sb))
        return -1;
    *out = hw * (sa / Sa + sb / Sb);
    return 0;
}

/* ------------------------------------------------------------------ */
/* pointwise kernels                                                   */
/* ------------------------------------------------------------------ */

/* weno_z.py:13-87 */
double or_weno_z_left(double v0, double v1, double v2, double v3, double v4)
{
    double p0 = v0 / 3.0 - 7.0 * v1 / 6.0 + 11.0 * v2 / 6.0;
    double p1 = -v1 / 6.0 + 5.0 * v2 / 6.0 + v3 / 3.0;
    double p2 = v2 / 3.0 + 5.0 * v3 / 6.0 - v4 / 6.0;
    double a = v0 - 2.0 * v1 + v2, b = v0 - 4.0 * v1 + 3.0 * v2;
    double beta0 = 13.0 / 12.0 * (a * a) + 0.25 * (b * b);
    a = v1 - 2.0 * v2 + v3;
    b = v1 - v3;
    double beta1 = 13.0 / 12.0 * (a * a) + 0.25 * (b * b);
    a = v2 - 2.0 * v3 + v4;
    b = 3.0 * v2 - 4.0 * v3 + v4;
    double beta2 = 13.0 / 12.0 * (a * a) + 0.25 * (b * b);
    double tau5 = fabs(beta2 - beta0);
    double r0 = tau5 / (beta0 + WENO_EPS);
    double r1 = tau5 / (beta1 + WENO_EPS);
    double r2 = tau5 / (beta2 + WENO_EPS);
    double alpha0 = 0.1 * (1.0 + r0 * r0);
    double alpha1 = 0.6 * (1.0 + r1 * r1);
    double alpha2 = 0.3 * (1.0 + r2 * r2);
    double s = alpha0 + alpha1 + alpha2;
    double w0 = alpha0 / s, w1 = alpha1 / s, w2 = alpha2 / s;
    return w0 * p0 + w1 * p1 + w2 * p2;
}

/* weno_z.py:91-129 / realizability.py:132-176 for one cell of a (4,n) array */
static void project_one(double *U, long n, long i, double eps)
{
    double phi = U[i];
    if (phi < 0.0)
        phi = 0.0;
    double jx = U[n + i], jy = U[2 * n + i], jz = U[3 * n + i];
    double jn = sqrt(jx * jx + jy * jy + jz * jz);
    double jmax = (1.0 - eps) * phi;
    if (jn > jmax) {
        double scale = jmax / jn;
        jx *= scale;
        jy *= scale;
        jz *= scale;
    }
    U[i] = phi;
    U[n + i] = jx;
    U[2 * n + i] = jy;
    U[3 * n + i] = jz;
}

void or_project_realizable(double *U, long n, double eps)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i)
        project_one(U, n, i, eps);
}

/* flux.py:63-135 / wave_speeds.py:37-103 */
static void eig_one(double phi, double jx, double jy, double jz, double jn,
                    int axis, double *lmin, double *lmax)
{
    double iso = 1.0 / sqrt(3.0);
    if (phi < EIG_EPS)
        phi = EIG_EPS;
    if (jn < EIG_EPS) {
        *lmin = -iso;
        *lmax = iso;
        return;
    }
    double f = jn / phi;
    if (f < 0.0)
        f = 0.0;
    else if (f > 1.0 - EIG_EPS)
        f = 1.0 - EIG_EPS;
    double jnorm = axis == 0 ? jx : (axis == 1 ? jy : jz);
    double mu = jnorm / jn;
    double chi = (5.0 - 2.0 * sqrt(4.0 - 3.0 * f * f)) / 3.0;
    double den = 5.0 - 3.0 * chi;
    double disc = 4.0 * f * f * mu * mu -
                  den * (mu * mu * (3.0 * chi - 1.0) - 2.0 * (1.0 - chi));
    if (disc < 0.0)
        disc = 0.0;
    double root = sqrt(disc);
    *lmin = (2.0 * f * mu - root) / den;
    *lmax = (2.0 * f * mu + root) / den;
}

void or_eigenvalues(const double *U, long n, int axis, double *lmin, double *lmax)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        double phi = U[i], jx = U[n + i], jy = U[2 * n + i], jz = U[3 * n + i];
        double jn = sqrt(jx * jx + jy * jy + jz * jz);
        eig_one(phi, jx, jy, jz, jn, axis, &lmin[i], &lmax[i]);
    }
}

/* flux.py:10-61 */
static void closure_one(double phi, double jx, double jy, double jz,
                        double *jn, double *nx, double *ny, double *nz,
                        double *iso, double *dir)
{
    double n = sqrt(jx * jx + jy * jy + jz * jz);
    double f;
    if (phi > 0.0) {
        f = n / phi;
        if (f < 0.0)
            f = 0.0;
        else if (f > 1.0)
            f = 1.0;
    } else {
        f = 0.0;
    }
    double chi = (5.0 - 2.0 * sqrt(4.0 - 3.0 * f * f)) / 3.0;
    if (n > 0.0) {
        *nx = jx / n;
        *ny = jy / n;
        *nz = jz / n;
    } else {
        *nx = *ny = *nz = 0.0;
    }
    *jn = n;
    *iso = 0.5 * (1.0 - chi);
    *dir = 0.5 * (3.0 * chi - 1.0);
}

/* physical flux of one state along axis (flux.py:199-309) */
static void phys_flux(int axis, double phi, double jx, double jy, double jz,
                      double nx, double ny, double nz, double iso, double dir,
                      double F[4])
{
    if (axis == 0) {
        F[0] = jx;
        F[1] = phi * (iso + dir * nx * nx);
        F[2] = phi * dir * nx * ny;
        F[3] = phi * dir * nx * nz;
    } else if (axis == 1) {
        F[0] = jy;
        F[1] = phi * dir * nx * ny;
        F[2] = phi * (iso + dir * ny * ny);
        F[3] = phi * dir * ny * nz;
    } else {
        F[0] = jz;
        F[1] = phi * dir * nx * nz;
        F[2] = phi * dir * ny * nz;
        F[3] = phi * (iso + dir * nz * nz);
    }
}

/* flux.py:137-386 */
void or_hll_flux(const double *UL, const double *UR, long n, int axis,
                 int wave_mode, double *F)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        double pl = UL[i], xl = UL[n + i], yl = UL[2 * n + i], zl = UL[3 * n + i];
        double pr = UR[i], xr = UR[n + i], yr = UR[2 * n + i], zr = UR[3 * n + i];
        double jnl, nxl, nyl, nzl, isol, dirl;
        double jnr, nxr, nyr, nzr, isor, dirr;
        closure_one(pl, xl, yl, zl, &jnl, &nxl, &nyl, &nzl, &isol, &dirl);
        closure_one(pr, xr, yr, zr, &jnr, &nxr, &nyr, &nzr, &isor, &dirr);
        double FL[4], FR[4];
        phys_flux(axis, pl, xl, yl, zl, nxl, nyl, nzl, isol, dirl, FL);
        phys_flux(axis, pr, xr, yr, zr, nxr, nyr, nzr, isor, dirr, FR);
        double sl, sr;
        if (wave_mode == OR_WAVE_EXACT) {
            double lminl, lmaxl, lminr, lmaxr;
            eig_one(pl, xl, yl, zl, jnl, axis, &lminl, &lmaxl);
            eig_one(pr, xr, yr, zr, jnr, axis, &lminr, &lmaxr);
            sl = lminl < lminr ? lminl : lminr;
            sr = lmaxl > lmaxr ? lmaxl : lmaxr;
            if (sl > 0.0)
                sl = 0.0;
            if (sr < 0.0)
                sr = 0.0;
        } else {
            sl = -1.0;
            sr = 1.0;
        }
        double den = sr - sl;
        double sp = sl * sr;
        F[i] = (sr * FL[0] - sl * FR[0] + sp * (pr - pl)) / den;
        F[n + i] = (sr * FL[1] - sl * FR[1] + sp * (xr - xl)) / den;
        F[2 * n + i] = (sr * FL[2] - sl * FR[2] + sp * (yr - yl)) / den;
        F[3 * n + i] = (sr * FL[3] - sl * FR[3] + sp * (zr - zl)) / den;
    }
}

/* ------------------------------------------------------------------ */
/* ghost cells, reconstruction, spatial operator                       */
/* ------------------------------------------------------------------ */

static inline long ix4(int c, long i, long j, long k, long nx, long ny, long nz)
{
    return (((long)c * nx + i) * ny + j) * nz + k;
}

/* boundary.py:117-399, ghost width 3 (boundary.py:247-320) */
int or_fill_ghost_cells(const double *U, int nx, int ny, int nz,
                        const double *U_in, double *Ue)
{
    if (nz < 4)
        return fail("WENO-Z inflow extrapolation requires at least four z cells.");
    const int g = GHOST;
    long ex = nx + 2 * g, ey = ny + 2 * g, ez = nz + 2 * g;
    /* interior + z-min quartic ghosts + z-max zero gradient, interior (x,y) */
#pragma omp parallel for collapse(2) schedule(static)
    for (int c = 0; c < 4; ++c) {
        for (int i = 0; i < nx; ++i) {
            for (int j = 0; j < ny; ++j) {
                const double *u = U + ix4(c, i, j, 0, nx, ny, nz);
                double *e = Ue + ix4(c, i + g, j + g, 0, ex, ey, ez);
                for (int k = 0; k < nz; ++k)
                    e[g + k] = u[k];
                double uin = U_in[((long)c * nx + i) * ny + j];
                double u0 = u[0], u1 = u[1], u2 = u[2], u3 = u[3];
                double G1 = 5.0 * uin - (77.0 / 12.0) * u0 + (43.0 / 12.0) * u1 -
                            (17.0 / 12.0) * u2 + 0.25 * u3;
                double G2 = 25.0 * uin - (505.0 / 12.0) * u0 + (335.0 / 12.0) * u1 -
                            (145.0 / 12.0) * u2 + (9.0 / 4.0) * u3;
                double G3 = 75.0 * uin - (545.0 / 4.0) * u0 + (395.0 / 4.0) * u1 -
                            (181.0 / 4.0) * u2 + (35.0 / 4.0) * u3;
                e[g - 1] = G1;
                e[g - 2] = G2;
                e[g - 3] = G3;
                for (int k = g + nz; k < ez; ++k)
                    e[k] = e[g + nz - 1];
            }
        }
    }
    /* x boundaries: copy plane g / g+nx-1 over all (y, z) -- boundary.py:347-369 */
    /* y boundaries: copy row g / g+ny-1 over all (x, z) -- boundary.py:375-397 */
    /* Equivalent closed form: U_ext(i,j,k) = U_ext(clamp_x(i), clamp_y(j), k). */
#pragma omp parallel for collapse(2) schedule(static)
    for (int c = 0; c < 4; ++c) {
        for (long i = 0; i < ex; ++i) {
            long si = i < g ? g : (i >= g + nx ? g + nx - 1 : i);
            for (long j = 0; j < ey; ++j) {
                long sj = j < g ? g : (j >= g + ny ? g + ny - 1 : j);
                if (si == i && sj == j)
                    continue;
                memcpy(Ue + ix4(c, i, j, 0, ex, ey, ez),
                       Ue + ix4(c, si, sj, 0, ex, ey, ez), sizeof(double) * (size_t)ez);
            }
        }
    }
    return 0;
}

/* weno_z.py:136-321 */
void or_reconstruct(const double *Ue, int nx, int ny, int nz, int axis,
                    double *UL, double *UR)
{
    const int g = GHOST;
    long ex = nx + 2 * g, ey = ny + 2 * g, ez = nz + 2 * g;
    long fx = nx + (axis == 0), fy = ny + (axis == 1), fz = nz + (axis == 2);
    long nf = fx * fy * fz;
    long stride = axis == 0 ? ey * ez : (axis == 1 ? ez : 1);
#pragma omp parallel for collapse(2) schedule(static)
    for (int c = 0; c < 4; ++c) {
        for (long i = 0; i < fx; ++i) {
            for (long j = 0; j < fy; ++j) {
                for (long k = 0; k < fz; ++k) {
                    /* first stencil cell in U_ext index space */
                    long bi = axis == 0 ? i : i + g;
                    long bj = axis == 1 ? j : j + g;
                    long bk = axis == 2 ? k : k + g;
                    const double *v = Ue + ix4(c, bi, bj, bk, ex, ey, ez);
                    long f = (i * fy + j) * fz + k;
                    UL[c * nf + f] = or_weno_z_left(v[0], v[stride], v[2 * stride],
                                                    v[3 * stride], v[4 * stride]);
                    UR[c * nf + f] = or_weno_z_left(v[5 * stride], v[4 * stride],
                                                    v[3 * stride], v[2 * stride],
                                                    v[stride]);
                }
            }
        }
    }
    or_project_realizable(UL, nf, REAL_EPS);
    or_project_realizable(UR, nf, REAL_EPS);
}

static int mat_check(const or_problem *p)
{
    for (long n = (long)p->nx * p->ny * p->nz, i = 0; i < n; ++i)
        if (p->ids[i] < 0 || p->ids[i] >= p->n_mat)
            return fail("material id out of range at cell %ld", i);
    return 0;
}

/* boundary.py:82-114 */
int or_inflow_state(const or_problem *p, double energy, double *Uin)
{
    if (p->energy_spread == 0)
        return fail("Zero energy spread is not supported by the Gaussian inflow.");
    double de = energy - p->beam_energy;
    double s = p->energy_spread;
    double ed = exp(-(de * de) / (2.0 * (s * s))) / (sqrt(2.0 * M_PI) * s);
    double rf = exp(-(p->angular_spread * p->angular_spread));
    long m = (long)p->nx * p->ny;
    for (long i = 0; i < m; ++i) {
        double phi = p->spatial[i] * ed;
        Uin[i] = phi;
        Uin[m + i] = 0.0;
        Uin[2 * m + i] = 0.0;
        Uin[3 * m + i] = rf * phi;
    }
    return 0;
}

int or_spatial_operator(const or_problem *p, int wave_mode, const double *U,
                        double energy, const double *bstate, double *rhs)
{
    int nx = p->nx, ny = p->ny, nz = p->nz;
    long n = (long)nx * ny * nz;
    long ne = (long)(nx + 6) * (ny + 6) * (nz + 6);
    long nfx = (long)(nx + 1) * ny * nz, nfy = (long)nx * (ny + 1) * nz,
         nfz = (long)nx * ny * (nz + 1);
    int rc = -1;
    double *Uin = NULL, *Ue = NULL, *L = NULL, *R = NULL, *Fx = NULL, *Fy = NULL,
           *Fz = NULL;
    Ue = malloc(sizeof(double) * 4 * (size_t)ne);
    long nfmax = nfx > nfy ? nfx : nfy;
    nfmax = nfmax > nfz ? nfmax : nfz;
    L = malloc(sizeof(double) * 4 * (size_t)nfmax);
    R = malloc(sizeof(double) * 4 * (size_t)nfmax);
    Fx = malloc(sizeof(double) * 4 * (size_t)nfx);
    Fy = malloc(sizeof(double) * 4 * (size_t)nfy);
    Fz = malloc(sizeof(double) * 4 * (size_t)nfz);
    if (!Ue || !L || !R || !Fx || !Fy || !Fz) {
        fail("out of memory");
        goto out;
    }
    if (!bstate) {
        Uin = malloc(sizeof(double) * 4 * (size_t)nx * ny);
        if (!Uin || or_inflow_state(p, energy, Uin))
            goto out;
        bstate = Uin;
    }
    if (or_fill_ghost_cells(U, nx, ny, nz, bstate, Ue))
        goto out;
    or_reconstruct(Ue, nx, ny, nz, 0, L, R);
    or_hll_flux(L, R, nfx, 0, wave_mode, Fx);
    or_reconstruct(Ue, nx, ny, nz, 1, L, R);
    or_hll_flux(L, R, nfy, 1, wave_mode, Fy);
    or_reconstruct(Ue, nx, ny, nz, 2, L, R);
    or_hll_flux(L, R, nfz, 2, wave_mode, Fz);
    /* spatial.py:239-243 */
    {
        double dx = p->dx, dy = p->dy, dz = p->dz;
#pragma omp parallel for collapse(2) schedule(static)
        for (int c = 0; c < 4; ++c) {
            for (long i = 0; i < nx; ++i) {
                for (long j = 0; j < ny; ++j) {
                    for (long k = 0; k < nz; ++k) {
                        double ax = (Fx[c * nfx + ((i + 1) * ny + j) * nz + k] -
                                     Fx[c * nfx + (i * ny + j) * nz + k]) / dx;
                        double ay = (Fy[c * nfy + (i * (ny + 1) + j + 1) * nz + k] -
                                     Fy[c * nfy + (i * (ny + 1) + j) * nz + k]) / dy;
                        double az = (Fz[c * nfz + (i * ny + j) * (nz + 1) + k + 1] -
                                     Fz[c * nfz + (i * ny + j) * (nz + 1) + k]) / dz;
                        rhs[c * n + (i * ny + j) * nz + k] = -(ax + ay + az);
                    }
                }
            }
        }
    }
    rc = 0;
out:
    free(Uin);
    free(Ue);
    free(L);
    free(R);
    free(Fx);
    free(Fy);
    free(Fz);
    return rc;
}

/* ------------------------------------------------------------------ */
/* per-material step scalars                                           */
/* ------------------------------------------------------------------ */

/* Per-material stopping powers at one energy (energy.py:40-57). */
static int S_per_mat(const or_problem *p, double energy, double *S)
{
    for (int m = 0; m < p->n_mat; ++m)
        if (or_stopping_power(&p->mats[m], energy, &S[m]))
            return -1;
    return 0;
}

/* energy.py:234-270, per material */
static int source_depths(const or_problem *p, const or_config *c, double energy,
                         double dE, double *ts, double *tn)
{
    for (int m = 0; m < p->n_mat; ++m) {
        ts[m] = 0.0;
        tn[m] = 0.0;
        if (dE <= 0.0)
            continue;
        if (c->scattering != OR_SCAT_OFF &&
            or_scattering_depth(&p->mats[m], c->scattering,
                                c->coefficient_interpolation, energy, dE, &ts[m]))
            return -1;
        if (c->nuclear && or_nuclear_depth(&p->mats[m], c->coefficient_interpolation,
                                           energy, dE, &tn[m]))
            return -1;
    }
    return 0;
}

/* energy.py:273-310 factors: f_phi = exp(sign*min(tn,700)),
 * f_J = exp(sign*min(tn+ts,700)) per material. */
static void source_factors(int n_mat, const double *ts, const double *tn,
                           int inverse, double *fphi, double *fj)
{
    double sign = inverse ? 1.0 : -1.0;
    for (int m = 0; m < n_mat; ++m) {
        double a = tn[m] < 700.0 ? tn[m] : 700.0;
        double s = tn[m] + ts[m];
        double b = s < 700.0 ? s : 700.0;
        fphi[m] = exp(sign * a);
        fj[m] = exp(sign * b);
    }
}

/* energy.py:83-104 written into B (4,nx,ny): S_b(E) * U_in(E) */
static int conserved_inflow(const or_problem *p, double energy, double *B)
{
    long m = (long)p->nx * p->ny;
    double S[64];
    if (S_per_mat(p, energy, S) || or_inflow_state(p, energy, B))
        return -1;
    for (long i = 0; i < m; ++i) {
        double s = S[p->ids[i * p->nz]];
        for (int c = 0; c < 4; ++c)
            B[c * m + i] = s * B[c * m + i];
    }
    return 0;
}

void or_rk3_trace_combination(long n, const double *B0, const double *B13,
                              const double *B23, const double *B1,
                              double *Bs1, double *Bs2)
{
    for (long i = 0; i < n; ++i) {
        Bs1[i] = 0.5 * (-9.0 * B0[i] + 18.0 * B13[i] - 9.0 * B23[i] + 2.0 * B1[i]);
        Bs2[i] = 0.25 * (11.0 * B0[i] - 27.0 * B13[i] + 27.0 * B23[i] - 7.0 * B1[i]);
    }
}

/* energy.py:107-230 and 348-455.  When ts/tn arrays are NULL the plain
 * SSPRK3 traces are produced; otherwise the IF-SSPRK3 traces with depths
 * for c = 1/3, 1/2, 2/3, 1 in that order (each n_mat long). */
static int stage_inflow(const or_problem *p, double energy, double dE,
                        const double *ts4[4], const double *tn4[4],
                        double *Bs1, double *Bs2)
{
    if (dE <= 0.0)
        return fail("dE must be positive.");
    long m = (long)p->nx * p->ny;
    int nm = p->n_mat;
    int rc = -1;
    double *B0 = malloc(sizeof(double) * 4 * (size_t)m * 4);
    if (!B0)
        return fail("out of memory");
    double *B13 = B0 + 4 * m, *B23 = B13 + 4 * m, *B1 = B23 + 4 * m;
    double Es[3] = {energy - dE / 3.0, energy - 2.0 * dE / 3.0, energy - dE};
    double *Bk[3] = {B13, B23, B1};
    const int depth_of[3] = {0, 2, 3}; /* c = 1/3, 2/3, 1 */
    if (conserved_inflow(p, energy, B0))
        goto out;
    for (int q = 0; q < 3; ++q) {
        if (conserved_inflow(p, Es[q], Bk[q]))
            goto out;
        if (ts4) {
            double fphi[64], fj[64];
            source_factors(nm, ts4[depth_of[q]], tn4[depth_of[q]], 1, fphi, fj);
            for (long i = 0; i < m; ++i) {
                int mat = p->ids[i * p->nz];
                Bk[q][i] *= fphi[mat];
                for (int c = 1; c < 4; ++c)
                    Bk[q][c * m + i] *= fj[mat];
            }
        }
    }
    or_rk3_trace_combination(4 * m, B0, B13, B23, B1, Bs1, Bs2);
    if (ts4) {
        double f1p[64], f1j[64], f2p[64], f2j[64];
        source_factors(nm, ts4[3], tn4[3], 0, f1p, f1j);
        source_factors(nm, ts4[1], tn4[1], 0, f2p, f2j);
        for (long i = 0; i < m; ++i) {
            int mat = p->ids[i * p->nz];
            Bs1[i] *= f1p[mat];
            Bs2[i] *= f2p[mat];
            for (int c = 1; c < 4; ++c) {
                Bs1[c * m + i] *= f1j[mat];
                Bs2[c * m + i] *= f2j[mat];
            }
        }
    }
    {
        double S1[64], S2[64];
        if (S_per_mat(p, energy - dE, S1) || S_per_mat(p, energy - 0.5 * dE, S2))
            goto out;
        for (long i = 0; i < m; ++i) {
            int mat = p->ids[i * p->nz];
            for (int c = 0; c < 4; ++c) {
                Bs1[c * m + i] = Bs1[c * m + i] / S1[mat];
                Bs2[c * m + i] = Bs2[c * m + i] / S2[mat];
            }
        }
    }
    rc = 0;
out:
    free(B0);
    return rc;
}

/* ------------------------------------------------------------------ */
/* CFL and integrators                                                 */
/* ------------------------------------------------------------------ */

/* energy.py:459-490: grow a support mask by `radius` cells along one axis,
 * one cell per pass. */
static void dilate_axis(const uint8_t *in, uint8_t *out, uint8_t *tmp, int nx,
                        int ny, int nz, int axis, int radius)
{
    long n = (long)nx * ny * nz;
    long stride = axis == 0 ? (long)ny * nz : axis == 1 ? nz : 1;
    int len = axis == 0 ? nx : axis == 1 ? ny : nz;
    memcpy(out, in, (size_t)n);
    for (int r = 0; r < radius; ++r) {
        memcpy(tmp, out, (size_t)n);
        for (long i = 0; i < n; ++i) {
            int pos = (int)((i / stride) % len);
            if (pos > 0 && tmp[i - stride])
                out[i] = 1;
            if (pos < len - 1 && tmp[i + stride])
                out[i] = 1;
        }
    }
}

/* energy.py:493-537: directional activity masks of the vacuum-aware CFL.
 * Uin is the (4, nx, ny) inflow state; masks is 3n bytes (x, y, z). */
static int vacuum_masks(const or_problem *p, const double *U, const double *Uin,
                        double tol, int radius, uint8_t *masks)
{
    int nx = p->nx, ny = p->ny, nz = p->nz;
    long n = (long)nx * ny * nz, m = (long)nx * ny;
    double mphi = U[0], min_ = Uin[0];
    for (long i = 1; i < n; ++i)
        if (U[i] > mphi)
            mphi = U[i];
    for (long i = 1; i < m; ++i)
        if (Uin[i] > min_)
            min_ = Uin[i];
    double scale = min_ > mphi ? min_ : mphi;
    if (scale == 0.0) {
        memset(masks, 1, (size_t)(3 * n));
        return 0;
    }
    double thr = tol * scale;
    uint8_t *act = malloc((size_t)n), *tmp = malloc((size_t)n);
    if (!act || !tmp) {
        free(act);
        free(tmp);
        return fail("out of memory");
    }
    for (long i = 0; i < n; ++i)
        act[i] = U[i] > thr;
    for (long col = 0; col < m; ++col)
        act[col * nz] |= Uin[col] > thr;
    for (int axis = 0; axis < 3; ++axis)
        dilate_axis(act, masks + axis * n, tmp, nx, ny, nz, axis, radius);
    free(act);
    free(tmp);
    return 0;
}

/* energy.py:539-660.  vac_tol < 0 stands for vacuum_relative_tolerance=None;
 * otherwise the directional speeds outside the WENO-Z stencil reach (3
 * cells) of the active cells are dropped. */
static int energy_step(const or_problem *p, const or_config *c, const double *U,
                       double energy, double vac_tol, double *dE_out)
{
    int nx = p->nx, ny = p->ny, nz = p->nz;
    long n = (long)nx * ny * nz, m = (long)nx * ny;
    double S[64];
    if (p->n_mat > 64)
        return fail("at most 64 materials");
    if (S_per_mat(p, energy, S))
        return -1;
    double mx = -INFINITY;
    if (c->wave_speeds == OR_WAVE_UNIT) {
        for (long i = 0; i < n; ++i) {
            double s = S[p->ids[i]];
            double v = 1.0 / (s * p->dx) + 1.0 / (s * p->dy) + 1.0 / (s * p->dz);
            if (v > mx)
                mx = v;
        }
    } else {
        double *lm = malloc(sizeof(double) * (size_t)n * 6);
        double *Uin = malloc(sizeof(double) * 4 * (size_t)m);
        double *li = malloc(sizeof(double) * 2 * (size_t)m);
        if (!lm || !Uin || !li) {
            free(lm);
            free(Uin);
            free(li);
            return fail("out of memory");
        }
        or_eigenvalues(U, n, 0, lm, lm + n);
        or_eigenvalues(U, n, 1, lm + 2 * n, lm + 3 * n);
        or_eigenvalues(U, n, 2, lm + 4 * n, lm + 5 * n);
        if (or_inflow_state(p, energy, Uin)) {
            free(lm);
            free(Uin);
            free(li);
            return -1;
        }
        or_eigenvalues(Uin, m, 2, li, li + m);
        uint8_t *mask = NULL;
        if (vac_tol >= 0.0) {
            mask = malloc((size_t)(3 * n));
            if (!mask || vacuum_masks(p, U, Uin, vac_tol, GHOST, mask)) {
                free(mask);
                free(lm);
                free(Uin);
                free(li);
                return mask ? -1 : fail("out of memory");
            }
        }
#pragma omp parallel for reduction(max : mx) schedule(static)
        for (long i = 0; i < n; ++i) {
            double ax = fmax(fabs(lm[i]), fabs(lm[n + i]));
            double ay = fmax(fabs(lm[2 * n + i]), fabs(lm[3 * n + i]));
            double az = fmax(fabs(lm[4 * n + i]), fabs(lm[5 * n + i]));
            if (i % nz == 0) {
                long col = i / nz;
                double ain = fmax(fabs(li[col]), fabs(li[m + col]));
                az = fmax(az, ain);
            }
            if (mask) {
                ax = mask[i] ? ax : 0.0;
                ay = mask[n + i] ? ay : 0.0;
                az = mask[2 * n + i] ? az : 0.0;
            }
            double s = S[p->ids[i]];
            double v = ax / (s * p->dx) + ay / (s * p->dy) + az / (s * p->dz);
            if (v > mx)
                mx = v;
        }
        free(mask);
        free(lm);
        free(Uin);
        free(li);
    }
    double dE = c->cfl / mx;
    *dE_out = dE <= c->max_energy_step ? dE : c->max_energy_step;
    return 0;
}

int or_compute_energy_step(const or_problem *p, const or_config *c,
                           const double *U, double energy, double *dE_out)
{
    return energy_step(p, c, U, energy, -1.0, dE_out);
}

int or_compute_energy_step_vacuum(const or_problem *p, const or_config *c,
                                  const double *U, double energy,
                                  double relative_tolerance, double *dE_out)
{
    if (!(relative_tolerance >= 0.0))
        return fail("relative tolerance must be non-negative");
    return energy_step(p, c, U, energy, relative_tolerance, dE_out);
}

/* energy.py:663-722 */
int or_forward_euler(const or_problem *p, const or_config *c, double *U,
                     double energy, double dE, double *edep)
{
    long n = (long)p->nx * p->ny * p->nz;
    double S0[64], S1[64];
    if (S_per_mat(p, energy, S0) || S_per_mat(p, energy - dE, S1))
        return -1;
    double *rhs = malloc(sizeof(double) * 4 * (size_t)n);
    if (!rhs)
        return fail("out of memory");
    if (or_spatial_operator(p, c->wave_speeds, U, energy, NULL, rhs)) {
        free(rhs);
        return -1;
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        int mt = p->ids[i];
        edep[i] = dE * U[i] * S0[mt];
        for (int q = 0; q < 4; ++q) {
            double w = rhs[q * n + i] * dE + S0[mt] * U[q * n + i];
            U[q * n + i] = w / S1[mt];
        }
    }
    or_project_realizable(U, n, REAL_EPS);
    free(rhs);
    return 0;
}

/* energy.py:725-873 (ts/tn NULL) and 878-1130 (IF form). */
static int rk3(const or_problem *p, const or_config *c, double *U, double energy,
               double dE, double *edep, int integrating_factor)
{
    long n = (long)p->nx * p->ny * p->nz, m = (long)p->nx * p->ny;
    int nm = p->n_mat;
    int rc = -1;
    double S0[64], S1[64], S2[64];
    double ts[4][64], tn[4][64];
    double f1p[64], f1j[64], f2p[64], f2j[64], i1p[64], i1j[64], i2p[64], i2j[64];
    const double *ts4[4], *tn4[4];
    if (integrating_factor) {
        /* energy.py:919-949: dE/3, 0.5*dE, 2*dE/3, dE */
        double dEs[4] = {dE / 3.0, 0.5 * dE, 2.0 * dE / 3.0, dE};
        for (int q = 0; q < 4; ++q) {
            if (source_depths(p, c, energy, dEs[q], ts[q], tn[q]))
                return -1;
            ts4[q] = ts[q];
            tn4[q] = tn[q];
        }
        source_factors(nm, ts[3], tn[3], 0, f1p, f1j);
        source_factors(nm, ts[1], tn[1], 0, f2p, f2j);
        source_factors(nm, ts[3], tn[3], 1, i1p, i1j);
        source_factors(nm, ts[1], tn[1], 1, i2p, i2j);
    }
    double *buf = malloc(sizeof(double) * (size_t)(4 * n * 4 + 8 * m + 3 * n));
    if (!buf)
        return fail("out of memory");
    double *W0 = buf, *V1 = W0 + 4 * n, *Us = V1 + 4 * n, *rhs = Us + 4 * n;
    double *Bs1 = rhs + 4 * n, *Bs2 = Bs1 + 4 * m;
    double *q0 = Bs2 + 4 * m, *q1 = q0 + n, *q2 = q1 + n;
    if (stage_inflow(p, energy, dE, integrating_factor ? ts4 : NULL,
                     integrating_factor ? tn4 : NULL, Bs1, Bs2))
        goto out;
    if (S_per_mat(p, energy, S0) || S_per_mat(p, energy - dE, S1) ||
        S_per_mat(p, energy - 0.5 * dE, S2))
        goto out;

    /* stage 0 */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        double s = S0[p->ids[i]];
        for (int q = 0; q < 4; ++q)
            W0[q * n + i] = s * U[q * n + i];
        q0[i] = U[i] * s;
    }
    if (or_spatial_operator(p, c->wave_speeds, U, energy, NULL, rhs))
        goto out;
    /* V1 = rhs*dE + V0; U1 = proj(C1 V1 / S1) */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        int mt = p->ids[i];
        for (int q = 0; q < 4; ++q) {
            double v = rhs[q * n + i] * dE + W0[q * n + i];
            V1[q * n + i] = v;
            double w = v;
            if (integrating_factor)
                w = v * (q == 0 ? f1p[mt] : f1j[mt]);
            Us[q * n + i] = w / S1[mt];
        }
    }
    or_project_realizable(Us, n, REAL_EPS);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i)
        q1[i] = Us[i] * S1[p->ids[i]];

    /* stage 1 */
    if (or_spatial_operator(p, c->wave_speeds, Us, energy - dE, Bs1, rhs))
        goto out;
    /* V2 = 0.75 V0 + 0.25 (C1^-1 rhs * dE + V1); U2 = proj(C2 V2 / S2).
     * V2 is written over V1 (not needed afterwards). */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        int mt = p->ids[i];
        for (int q = 0; q < 4; ++q) {
            double r = rhs[q * n + i];
            if (integrating_factor)
                r = r * (q == 0 ? i1p[mt] : i1j[mt]);
            r = r * dE;
            r = r + V1[q * n + i];
            r = r * 0.25;
            double v = 0.75 * W0[q * n + i] + r;
            V1[q * n + i] = v;
            double w = v;
            if (integrating_factor)
                w = v * (q == 0 ? f2p[mt] : f2j[mt]);
            Us[q * n + i] = w / S2[mt];
        }
    }
    or_project_realizable(Us, n, REAL_EPS);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i)
        q2[i] = Us[i] * S2[p->ids[i]];

    /* stage 2 */
    if (or_spatial_operator(p, c->wave_speeds, Us, energy - 0.5 * dE, Bs2, rhs))
        goto out;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        int mt = p->ids[i];
        for (int q = 0; q < 4; ++q) {
            double r = rhs[q * n + i];
            if (integrating_factor)
                r = r * (q == 0 ? i2p[mt] : i2j[mt]);
            r = r * dE;
            r = r + V1[q * n + i];
            r = r * (2.0 / 3.0);
            double v = (1.0 / 3.0) * W0[q * n + i] + r;
            if (integrating_factor)
                v = v * (q == 0 ? f1p[mt] : f1j[mt]);
            U[q * n + i] = v / S1[mt];
        }
        edep[i] = dE * ((1.0 / 6.0) * q0[i] + (1.0 / 6.0) * q1[i] +
                        (2.0 / 3.0) * q2[i]);
    }
    or_project_realizable(U, n, REAL_EPS);
    rc = 0;
out:
    free(buf);
    return rc;
}

int or_ssprk3(const or_problem *p, const or_config *c, double *U, double energy,
              double dE, double *edep)
{
    return rk3(p, c, U, energy, dE, edep, 0);
}

/* energy.py:878-1130 */
int or_if_ssprk3(const or_problem *p, const or_config *c, double *U,
                 double energy, double dE, double *edep)
{
    if (c->scattering == OR_SCAT_OFF && !c->nuclear)
        return rk3(p, c, U, energy, dE, edep, 0);
    return rk3(p, c, U, energy, dE, edep, 1);
}

/* energy.py:1132-1165 */
static int advance(const or_problem *p, const or_config *c, double *U,
                   double energy, double dE, double *edep)
{
    if (c->energy_scheme == OR_FORWARD_EULER)
        return or_forward_euler(p, c, U, energy, dE, edep);
    if (c->energy_scheme == OR_SSPRK3 || c->energy_scheme == OR_IF_SSPRK3)
        return or_ssprk3(p, c, U, energy, dE, edep);
    return fail("Energy scheme has not been migrated yet.");
}

/* energy.py:1168-1191: half-step factors per material */
static void half_source(const or_problem *p, const or_config *c, double *U,
                        const double *eta, const double *taun)
{
    long n = (long)p->nx * p->ny * p->nz;
    double fn[64], fs[64];
    for (int m = 0; m < p->n_mat; ++m) {
        double a = 0.5 * taun[m], b = 0.5 * eta[m];
        fn[m] = exp(-(a < 700.0 ? a : 700.0));
        fs[m] = exp(-(b < 700.0 ? b : 700.0));
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        int mt = p->ids[i];
        if (c->nuclear)
            for (int q = 0; q < 4; ++q)
                U[q * n + i] *= fn[mt];
        if (c->scattering != OR_SCAT_OFF)
            for (int q = 1; q < 4; ++q)
                U[q * n + i] *= fs[mt];
    }
}

/* energy.py:1194-1260 */
int or_strang_step(const or_problem *p, const or_config *c, double *U,
                   double energy, double dE, double *edep)
{
    double eta[64], taun[64];
    for (int m = 0; m < p->n_mat; ++m) {
        eta[m] = 0.0;
        taun[m] = 0.0;
        if (c->scattering != OR_SCAT_OFF &&
            or_scattering_depth(&p->mats[m], c->scattering,
                                c->coefficient_interpolation, energy, dE, &eta[m]))
            return -1;
        if (c->nuclear && or_nuclear_depth(&p->mats[m], c->coefficient_interpolation,
                                           energy, dE, &taun[m]))
            return -1;
    }
    half_source(p, c, U, eta, taun);
    if (advance(p, c, U, energy, dE, edep))
        return -1;
    half_source(p, c, U, eta, taun);
    return 0;
}

static int cmp_desc(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x < y) - (x > y);
}

/* energy.py:1497-1644 */
int or_energy_march(const or_problem *p, const or_config *c, double *U,
                    double *dose, double *energies, int cap, int *n_energies,
                    const double *snap_E, int n_snap, double *snap_out,
                    int *snap_taken, int max_steps)
{
    long n = (long)p->nx * p->ny * p->nz;
    if (p->n_mat > 64 || p->n_mat < 1)
        return fail("between 1 and 64 materials are supported");
    if (mat_check(p))
        return -1;
    double energy = p->beam_energy + c->energy_window_sigma * p->energy_spread;
    double emin = c->energy_min;
    /* targets: unique requested energies inside [emin, energy], descending */
    double *targets = malloc(sizeof(double) * (size_t)(n_snap > 0 ? n_snap : 1));
    int *tmap = malloc(sizeof(int) * (size_t)(n_snap > 0 ? n_snap : 1));
    double *edep = malloc(sizeof(double) * (size_t)n);
    if (!targets || !tmap || !edep) {
        free(targets);
        free(tmap);
        free(edep);
        return fail("out of memory");
    }
    int nt = 0;
    for (int i = 0; i < n_snap; ++i) {
        if (snap_taken)
            snap_taken[i] = 0;
        double t = snap_E[i];
        if (!(emin <= t && t <= energy))
            continue;
        int dup = 0;
        for (int k = 0; k < nt; ++k)
            dup |= targets[k] == t;
        if (!dup)
            targets[nt++] = t;
    }
    qsort(targets, (size_t)nt, sizeof(double), cmp_desc);
    for (int k = 0; k < nt; ++k)
        for (int i = 0; i < n_snap; ++i)
            if (snap_E[i] == targets[k]) {
                tmap[k] = i;
                break;
            }
    double rho[64];
    for (int m = 0; m < p->n_mat; ++m)
        rho[m] = p->mats[m].density;
    memset(dose, 0, sizeof(double) * (size_t)n);
    int ne = 0, ti = 0, steps = 0, rc = -1;
    if (ne < cap)
        energies[ne] = energy;
    ne++;
    while (energy > emin) {
        if (max_steps > 0 && steps >= max_steps)
            break;
        double dE;
        if (or_compute_energy_step(p, c, U, energy, &dE))
            goto out;
        if (ti < nt) {
            double t = targets[ti];
            if (energy > t && t >= energy - dE)
                dE = energy - t;
        }
        if (energy - emin < dE)
            dE = energy - emin;
        int r;
        if (c->energy_scheme == OR_IF_SSPRK3)
            r = or_if_ssprk3(p, c, U, energy, dE, edep);
        else if (c->scattering != OR_SCAT_OFF || c->nuclear)
            r = or_strang_step(p, c, U, energy, dE, edep);
        else
            r = advance(p, c, U, energy, dE, edep);
        if (r)
            goto out;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i)
            dose[i] += edep[i] * MEV_TO_JOULE / rho[p->ids[i]];
        energy -= dE;
        if (ne < cap)
            energies[ne] = energy;
        ne++;
        steps++;
        if (ti < nt) {
            double t = targets[ti];
            if (fabs(energy - t) <= 1e-12) {
                int slot = tmap[ti];
                if (snap_out)
                    memcpy(snap_out + (size_t)slot * 4 * (size_t)n, U,
                           sizeof(double) * 4 * (size_t)n);
                if (snap_taken)
                    snap_taken[slot] = 1;
                ti++;
            }
        }
    }
    rc = 0;
out:
    *n_energies = ne;
    free(targets);
    free(tmap);
    free(edep);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Moving active domain                                                */
/* ------------------------------------------------------------------ */

/* moving_active_domain.py:35-60: beam support + guard, z in [0, 1 + guard). */
static int initial_box(const or_problem *p, int guard, int b[6])
{
    int ix0 = -1, ix1 = -1, iy0 = -1, iy1 = -1;
    for (int i = 0; i < p->nx; ++i)
        for (int j = 0; j < p->ny; ++j)
            if (p->spatial[(long)i * p->ny + j] > 0.0) {
                if (ix0 < 0)
                    ix0 = i;
                ix1 = i;
                if (iy0 < 0 || j < iy0)
                    iy0 = j;
                if (j > iy1)
                    iy1 = j;
            }
    if (ix0 < 0)
        return fail("Beam inflow has empty transverse support.");
    b[0] = ix0 - guard > 0 ? ix0 - guard : 0;
    b[1] = ix1 + 1 + guard < p->nx ? ix1 + 1 + guard : p->nx;
    b[2] = iy0 - guard > 0 ? iy0 - guard : 0;
    b[3] = iy1 + 1 + guard < p->ny ? iy1 + 1 + guard : p->ny;
    b[4] = 0;
    b[5] = 1 + guard < p->nz ? 1 + guard : p->nz;
    return 0;
}

/* moving_active_domain.py:63-104: the box's grid, material ids and inflow
 * (the beam profile while the box touches z = 0, zero otherwise).  ids and
 * spatial are caller buffers of the full grid's sizes. */
static void local_problem(const or_problem *p, const int b[6], int32_t *ids,
                          double *spatial, or_problem *lp)
{
    int nx = b[1] - b[0], ny = b[3] - b[2], nz = b[5] - b[4];
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            long g = ((long)(b[0] + i) * p->ny + (b[2] + j)) * p->nz + b[4];
            long l = ((long)i * ny + j) * nz;
            memcpy(ids + l, p->ids + g, sizeof(int32_t) * (size_t)nz);
            spatial[(long)i * ny + j] =
                b[4] == 0 ? p->spatial[(long)(b[0] + i) * p->ny + (b[2] + j)] : 0.0;
        }
    *lp = *p;
    lp->nx = nx;
    lp->ny = ny;
    lp->nz = nz;
    lp->ids = ids;
    lp->spatial = spatial;
}

/* moving_active_domain.py:107-151: bounding box of phi > tol * scale (scale:
 * the largest phi, and of the inflow phi while k0 == 0).  Returns 1 and the
 * global box, or 0 for "no significant cell" (None). */
static int significant_box(const or_problem *lp, const int b[6], const double *U,
                           double energy, double tol, double *Uin, int out[6])
{
    long n = (long)lp->nx * lp->ny * lp->nz, m = (long)lp->nx * lp->ny;
    int inflow = b[4] == 0;
    double scale = U[0];
    for (long i = 1; i < n; ++i)
        if (U[i] > scale)
            scale = U[i];
    if (inflow) {
        if (or_inflow_state(lp, energy, Uin))
            return -1;
        double mi = Uin[0];
        for (long c = 1; c < m; ++c)
            if (Uin[c] > mi)
                mi = Uin[c];
        if (mi > scale)
            scale = mi;
    }
    if (scale == 0.0)
        return 0;
    double thr = tol * scale;
    int lo[3] = {lp->nx, lp->ny, lp->nz}, hi[3] = {-1, -1, -1};
    for (int i = 0; i < lp->nx; ++i)
        for (int j = 0; j < lp->ny; ++j)
            for (int k = 0; k < lp->nz; ++k) {
                long l = ((long)i * lp->ny + j) * lp->nz + k;
                int act = U[l] > thr;
                if (inflow && k == 0)
                    act |= Uin[(long)i * lp->ny + j] > thr;
                if (!act)
                    continue;
                int q[3] = {i, j, k};
                for (int a = 0; a < 3; ++a) {
                    if (q[a] < lo[a])
                        lo[a] = q[a];
                    if (q[a] > hi[a])
                        hi[a] = q[a];
                }
            }
    if (hi[0] < 0)
        return 0;
    for (int a = 0; a < 3; ++a) {
        out[2 * a] = b[2 * a] + lo[a];
        out[2 * a + 1] = b[2 * a] + hi[a] + 1;
    }
    return 1;
}

/* moving_active_domain.py:196-232 (also the crop of a full-grid state and
 * embed_state, 235-244, as boxes of the whole grid): dst on box nb holds src
 * on box ob where they overlap and zero elsewhere. */
static void rebox(const double *src, const int ob[6], double *dst, const int nb[6])
{
    int onx = ob[1] - ob[0], ony = ob[3] - ob[2], onz = ob[5] - ob[4];
    int nnx = nb[1] - nb[0], nny = nb[3] - nb[2], nnz = nb[5] - nb[4];
    long on = (long)onx * ony * onz, nn = (long)nnx * nny * nnz;
    memset(dst, 0, sizeof(double) * 4 * (size_t)nn);
    int g0[3], g1[3];
    for (int a = 0; a < 3; ++a) {
        g0[a] = ob[2 * a] > nb[2 * a] ? ob[2 * a] : nb[2 * a];
        g1[a] = ob[2 * a + 1] < nb[2 * a + 1] ? ob[2 * a + 1] : nb[2 * a + 1];
        if (g1[a] <= g0[a])
            return;
    }
    for (int c = 0; c < 4; ++c)
        for (int gi = g0[0]; gi < g1[0]; ++gi)
            for (int gj = g0[1]; gj < g1[1]; ++gj) {
                long so = c * on + ((long)(gi - ob[0]) * ony + (gj - ob[2])) * onz + (g0[2] - ob[4]);
                long d = c * nn + ((long)(gi - nb[0]) * nny + (gj - nb[2])) * nnz + (g0[2] - nb[4]);
                memcpy(dst + d, src + so, sizeof(double) * (size_t)(g1[2] - g0[2]));
            }
}

/* energy.py:1262-1494 with moving_active_domain.py:35-253. */
int or_energy_march_moving(const or_problem *p, const or_config *c, int guard,
                           double truncation_sigma, int vacuum_aware,
                           const double *U0, double *U_out, double *dose,
                           double *energies, int cap, int *n_energies,
                           const double *snap_E, int n_snap, double *snap_out,
                           int *snap_taken, int max_steps, int32_t *boxes)
{
    long n = (long)p->nx * p->ny * p->nz, m = (long)p->nx * p->ny;
    int full[6] = {0, p->nx, 0, p->ny, 0, p->nz};
    *n_energies = 0;
    if (p->n_mat > 64 || p->n_mat < 1)
        return fail("between 1 and 64 materials are supported");
    if (guard < 4)
        return fail("active_guard_cells must be an integer >= 4.");
    if (!(truncation_sigma > 0.0))
        return fail("active_truncation_sigma must be positive.");
    if (mat_check(p))
        return -1;
    double energy = p->beam_energy + c->energy_window_sigma * p->energy_spread;
    double emin = c->energy_min;
    double tol = exp(-0.5 * (truncation_sigma * truncation_sigma));
    int ns = n_snap > 0 ? n_snap : 1;
    double *targets = malloc(sizeof(double) * (size_t)ns);
    int *tmap = malloc(sizeof(int) * (size_t)ns);
    double *U = malloc(sizeof(double) * 4 * (size_t)n);
    double *W = malloc(sizeof(double) * 4 * (size_t)n);
    double *edep = malloc(sizeof(double) * (size_t)n);
    int32_t *ids = malloc(sizeof(int32_t) * (size_t)n);
    double *spatial = malloc(sizeof(double) * (size_t)m);
    double *Uin = malloc(sizeof(double) * 4 * (size_t)m);
    int rc = -1;
    if (!targets || !tmap || !U || !W || !edep || !ids || !spatial || !Uin) {
        fail("out of memory");
        goto out;
    }
    int nt = 0;
    for (int i = 0; i < n_snap; ++i) {
        if (snap_taken)
            snap_taken[i] = 0;
        double t = snap_E[i];
        if (!(emin <= t && t <= energy))
            continue;
        int dup = 0;
        for (int k = 0; k < nt; ++k)
            dup |= targets[k] == t;
        if (!dup)
            targets[nt++] = t;
    }
    qsort(targets, (size_t)nt, sizeof(double), cmp_desc);
    for (int k = 0; k < nt; ++k)
        for (int i = 0; i < n_snap; ++i)
            if (snap_E[i] == targets[k]) {
                tmap[k] = i;
                break;
            }
    double rho[64];
    for (int q = 0; q < p->n_mat; ++q)
        rho[q] = p->mats[q].density;
    int box[6];
    if (initial_box(p, guard, box))
        goto out;
    if (U0)
        rebox(U0, full, U, box);
    else
        memset(U, 0, sizeof(double) * 4 * (size_t)n);
    memset(dose, 0, sizeof(double) * (size_t)n);
    int ne = 0, ti = 0, steps = 0;
    if (ne < cap)
        energies[ne] = energy;
    ne++;
    or_problem lp;
    while (energy > emin) {
        if (max_steps > 0 && steps >= max_steps)
            break;
        local_problem(p, box, ids, spatial, &lp);
        int sig[6], next[6];
        int r = significant_box(&lp, box, U, energy, tol, Uin, sig);
        if (r < 0)
            goto out;
        if (r == 0) {
            if (initial_box(p, guard, next))
                goto out;
        } else {
            next[0] = sig[0] - guard > 0 ? sig[0] - guard : 0;
            next[1] = sig[1] + guard < p->nx ? sig[1] + guard : p->nx;
            next[2] = sig[2] - guard > 0 ? sig[2] - guard : 0;
            next[3] = sig[3] + guard < p->ny ? sig[3] + guard : p->ny;
            next[4] = sig[4] - guard > 0 ? sig[4] - guard : 0;
            next[5] = sig[5] + guard < p->nz ? sig[5] + guard : p->nz;
        }
        if (memcmp(next, box, sizeof box) != 0) {
            rebox(U, box, W, next);
            double *t = U;
            U = W;
            W = t;
            memcpy(box, next, sizeof box);
            local_problem(p, box, ids, spatial, &lp);
        }
        if (boxes && steps < cap)
            memcpy(boxes + 6 * (size_t)steps, box, sizeof box);
        double dE;
        if (energy_step(&lp, c, U, energy, vacuum_aware ? tol : -1.0, &dE))
            goto out;
        if (ti < nt) {
            double t = targets[ti];
            if (energy > t && t >= energy - dE)
                dE = energy - t;
        }
        if (energy - emin < dE)
            dE = energy - emin;
        if (c->energy_scheme == OR_IF_SSPRK3)
            r = or_if_ssprk3(&lp, c, U, energy, dE, edep);
        else if (c->scattering != OR_SCAT_OFF || c->nuclear)
            r = or_strang_step(&lp, c, U, energy, dE, edep);
        else
            r = advance(&lp, c, U, energy, dE, edep);
        if (r)
            goto out;
        for (int i = 0; i < lp.nx; ++i)
            for (int j = 0; j < lp.ny; ++j)
                for (int k = 0; k < lp.nz; ++k) {
                    long l = ((long)i * lp.ny + j) * lp.nz + k;
                    long g = ((long)(box[0] + i) * p->ny + (box[2] + j)) * p->nz + (box[4] + k);
                    dose[g] += edep[l] * MEV_TO_JOULE / rho[ids[l]];
                }
        energy -= dE;
        if (ne < cap)
            energies[ne] = energy;
        ne++;
        steps++;
        if (ti < nt) {
            double t = targets[ti];
            if (fabs(energy - t) <= 1e-12) {
                int slot = tmap[ti];
                if (snap_out)
                    rebox(U, box, snap_out + (size_t)slot * 4 * (size_t)n, full);
                if (snap_taken)
                    snap_taken[slot] = 1;
                ti++;
            }
        }
    }
    if (U_out)
        rebox(U, box, U_out, full);
    *n_energies = ne;
    rc = 0;
out:
    free(targets);
    free(tmap);
    free(U);
    free(W);
    free(edep);
    free(ids);
    free(spatial);
    free(Uin);
    return rc;
}
