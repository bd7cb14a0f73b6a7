"""Dose metrics used to state the accuracy of every march result.

Restates the deterministic-dose parts of ``proton_solver/metrics.py``:
maximum dose and its relative difference (metrics.py:107-112, 239-253), the
central-axis cell and Bragg-peak depth (146-235), the laterally integrated
depth dose with R80 and distal fall-off width (331-460), plus a global 3-D
gamma index (metrics.py:542-606 semantics: global normalisation to the
reference maximum, lower dose cutoff, distance-to-agreement search).  The
reference computes gamma with pymedphys; this module uses its own exhaustive
search over a sub-voxel grid, so its pass rates are not bit-compared with
pymedphys.
"""

from __future__ import annotations

import math

import numpy as np


def maximum_dose_relative_difference(dose, reference) -> float:
    ref = float(np.max(reference))
    return (float(np.max(dose)) - ref) / ref


def max_abs_difference_over_max(dose, reference) -> float:
    """max|D - D_ref| / max D_ref, the FP64 parity measure."""
    ref = float(np.max(reference))
    return float(np.max(np.abs(np.asarray(dose) - np.asarray(reference)))) / ref


def axis_cell_index(centres) -> int:
    """Cell containing coordinate 0 with half-open cells (metrics.py:146-171)."""
    centres = np.asarray(centres, dtype=float)
    if centres.size == 1:
        return 0
    width = float(centres[1] - centres[0])
    first_lower_face = float(centres[0]) - 0.5 * width
    index = int(np.floor((0.0 - first_lower_face) / width + 1e-9))
    return min(max(index, 0), centres.size - 1)


def central_axis_dose(dose, x, y):
    return np.asarray(dose)[axis_cell_index(x), axis_cell_index(y), :]


def peak_depth(dose, x, y, z) -> float:
    return float(np.asarray(z)[int(np.argmax(central_axis_dose(dose, x, y)))])


def integrated_depth_dose(dose):
    return np.asarray(dose, dtype=float).sum(axis=(0, 1))


def distal_peak_index(curve, window=10, min_fraction=0.3) -> int:
    """Deepest local peak above min_fraction of the maximum (metrics.py:340-362)."""
    curve = np.asarray(curve, dtype=float)
    threshold = min_fraction * float(np.max(curve))
    n = len(curve)
    for i in range(n - 1, -1, -1):
        if curve[i] < threshold:
            continue
        lo, hi = max(0, i - window), min(n, i + window + 1)
        if curve[i] >= float(np.max(curve[lo:hi])):
            return i
    return int(np.argmax(curve))


def distal_depth(z, curve, level, z_from=None):
    """Depth where the curve falls to level*peak beyond the peak (metrics.py:364-391)."""
    z = np.asarray(z, dtype=float)
    curve = np.asarray(curve, dtype=float)
    i_start = 0 if z_from is None else int(np.searchsorted(z, z_from))
    i_start = min(max(i_start, 0), len(curve) - 1)
    i_peak = i_start + int(np.argmax(curve[i_start:]))
    target = level * float(curve[i_peak])
    for i in range(i_peak, len(curve) - 1):
        if curve[i] >= target > curve[i + 1]:
            return float(z[i] + (target - curve[i]) * (z[i + 1] - z[i]) / (curve[i + 1] - curve[i]))
    return None


def range_metrics(dose, z, z_from=None):
    """(R80, R20 - R80) in metres of the integrated depth dose (metrics.py:399-417)."""
    idd = integrated_depth_dose(dose)
    r80 = distal_depth(z, idd, 0.8, z_from)
    r20 = distal_depth(z, idd, 0.2, z_from)
    if r80 is None or r20 is None:
        return r80, None
    return r80, r20 - r80


def _gamma_setup(dose, reference, spacing_m, dose_percent, distance_mm, lower_percent_cutoff,
                 interp_fraction, max_gamma):
    """Evaluated points, search offsets (in voxels, sorted by length), their
    distance terms and the dose tolerance of gamma_pass_rate; None when no
    point is evaluated."""
    dose = np.asarray(dose, dtype=float)
    ref = np.asarray(reference, dtype=float)
    if dose.shape != ref.shape:
        raise ValueError("dose and reference must have the same shape")
    dmax = float(ref.max())
    if dmax <= 0.0:
        return None
    dd = dose_percent / 100.0 * dmax
    dist = distance_mm / 1000.0
    mask = ref >= lower_percent_cutoff / 100.0 * dmax
    pts = np.argwhere(mask)
    if len(pts) == 0:
        return None
    step = dist / interp_fraction
    reach = max_gamma * dist
    offs = []
    n = int(math.ceil(reach / step))
    for a in range(-n, n + 1):
        for b in range(-n, n + 1):
            for c in range(-n, n + 1):
                r = math.sqrt((a * step) ** 2 + (b * step) ** 2 + (c * step) ** 2)
                if r <= reach + 1e-15:
                    offs.append((a * step, b * step, c * step, r))
    offs = np.asarray(offs)
    offs = offs[np.argsort(offs[:, 3], kind="stable")]
    scaled = offs[:, :3] / np.asarray(spacing_m, dtype=float)
    dist_terms = np.array([(r / dist) ** 2 for r in offs[:, 3]])
    return dose, ref, pts, scaled, dist_terms, dd


def gamma_pass_rate(dose, reference, spacing_m, dose_percent=1.0, distance_mm=1.0,
                    lower_percent_cutoff=10.0, interp_fraction=10, max_gamma=1.1,
                    device=None):
    """Global 3-D gamma pass rate in percent.

    Evaluated points are reference voxels at or above ``lower_percent_cutoff``
    of the reference maximum.  For each, the evaluated dose is searched on a
    grid of offsets with step ``distance/interp_fraction`` inside a sphere of
    radius ``max_gamma * distance``; the evaluated dose between voxel centres is
    trilinearly interpolated; search points outside the grid do not match.
    gamma = min over offsets of
    sqrt((dD / (dose_percent% * max_ref))^2 + (r / distance)^2); a point passes
    when gamma <= 1.

    ``device`` (a HIP device index) evaluates the points on the GPU
    (m1_gamma_pass_count in libm1march.so), bit-identical to the host search.
    """
    setup = _gamma_setup(dose, reference, spacing_m, dose_percent, distance_mm,
                         lower_percent_cutoff, interp_fraction, max_gamma)
    if setup is None:
        return None
    dose, ref, pts, scaled, dist_terms, dd = setup
    if device is not None:
        from . import _native

        return 100.0 * _native.gamma_pass_count(device, dose, ref, pts, scaled, dist_terms,
                                                dd) / len(pts)
    shape = np.asarray(dose.shape)
    gamma2 = np.full(len(pts), np.inf)
    ref_vals = ref[tuple(pts.T)]
    for k in range(len(scaled)):
        dist_term = dist_terms[k]
        alive = gamma2 > dist_term
        if not np.any(alive):
            break
        p = pts[alive].astype(float) + scaled[k]
        # a search point outside the grid never matches (pymedphys
        # interpolates the evaluated dose with an infinite fill value there)
        inside = np.all((p >= 0.0) & (p <= shape - 1.0), axis=1)
        p = np.clip(p, 0.0, shape - 1.0)
        i0 = np.floor(p).astype(int)
        i0 = np.minimum(i0, shape - 2).clip(min=0)
        f = p - i0
        val = np.zeros(len(p))
        for dx_ in (0, 1):
            wx = f[:, 0] if dx_ else 1.0 - f[:, 0]
            ix = np.minimum(i0[:, 0] + dx_, shape[0] - 1)
            for dy_ in (0, 1):
                wy = f[:, 1] if dy_ else 1.0 - f[:, 1]
                iy = np.minimum(i0[:, 1] + dy_, shape[1] - 1)
                for dz_ in (0, 1):
                    wz = f[:, 2] if dz_ else 1.0 - f[:, 2]
                    iz = np.minimum(i0[:, 2] + dz_, shape[2] - 1)
                    val += wx * wy * wz * dose[ix, iy, iz]
        g2 = np.where(inside, ((val - ref_vals[alive]) / dd) ** 2 + dist_term, np.inf)
        cur = gamma2[alive]
        gamma2[alive] = np.minimum(cur, g2)
    return 100.0 * float(np.mean(gamma2 <= 1.0))
