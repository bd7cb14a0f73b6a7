"""Shared helpers of the test suite (configs from fixtures, paths)."""

import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "proton-dose-m1_amd"
for p in (str(PKG), str(ROOT)):
    if p not in sys.path:
        sys.path.insert(0, p)

from proton_solver import (  # noqa: E402
    Beam, Box, CoefficientInterpolation, EnergyScheme, Geometry, Material,
    Numerics, Physics, Scattering, SolverConfig, SpatialScheme, WaveSpeeds,
)

GOLDEN = ROOT / "tests" / "golden"


def thesis_cases():
    return json.loads((GOLDEN / "thesis_golden.json").read_text())["cases"]


def config_from_fixture(c: dict) -> SolverConfig:
    """SolverConfig of a thesis run recorded in thesis_golden.json."""
    g = c["geometry"]
    objects = [
        Box(o["x_min"], o["x_max"], o["y_min"], o["y_max"], o["z_min"], o["z_max"],
            Material[o["material"]])
        for o in g["objects"]
    ]
    b = c["beam"]
    n = c["numerics"]
    p = c["physics"]
    return SolverConfig(
        geometry=Geometry(size=tuple(g["size"]), resolution=tuple(g["resolution"]),
                          background_material=Material[g["background_material"]],
                          objects=objects),
        beam=Beam(n_protons=b["n_protons"], energy=b["energy"], energy_spread=b["energy_spread"],
                  sigma=tuple(b["sigma"]), center=tuple(b["center"]),
                  angular_spread=b["angular_spread"], cutoff_sigma=b["cutoff_sigma"]),
        physics=Physics(scattering=Scattering[p["scattering"]], nuclear=p["nuclear"],
                        coefficient_interpolation=CoefficientInterpolation[
                            p.get("coefficient_interpolation", "LINEAR")]),
        numerics=Numerics(spatial_scheme=SpatialScheme[n["spatial_scheme"]],
                          energy_scheme=EnergyScheme[n["energy_scheme"]],
                          wave_speeds=WaveSpeeds[n["wave_speeds"]], cfl=n["cfl"],
                          energy_min=n["energy_min"], energy_window_sigma=n["energy_window_sigma"],
                          max_energy_step=n["max_energy_step"],
                          **{k: n[k] for k in ("moving_active_domain", "active_guard_cells",
                                               "active_truncation_sigma", "vacuum_aware_cfl")
                             if k in n}),
    )


def ct_golden():
    return json.loads((GOLDEN / "ct_golden.json").read_text())


def ct_section_npz(directory, upsample=6):
    """The reference CT benchmark's section (tests/golden/ct_section.npz),
    refined `upsample` times by voxel repetition, written as a CT NPZ."""
    import numpy as np

    with np.load(GOLDEN / "ct_section.npz", allow_pickle=False) as d:
        hu, body = d["HU"], d["body_mask"]
    for axis in range(3):
        hu = np.repeat(hu, upsample, axis)
        body = np.repeat(body, upsample, axis)
    path = Path(directory) / "ct_section_refined.npz"
    np.savez(path, HU=hu, body_mask=body)
    return path


def ct_config(path, upsample=6) -> SolverConfig:
    """SolverConfig of the reference CT benchmark
    (thesis_experiments/scattering_nuclear/ct.py: 150 MeV, Wentzel + nuclear,
    IF-SSPRK3, water / compact bone split at 150 HU)."""
    from proton_solver import HUPartition

    h = 3.0e-3 / upsample
    geometry = Geometry.from_ct_scan(
        path, source_spacing_m=(h, h, h),
        hu_partition=HUPartition(cutoffs=(150.0,),
                                 materials=(Material.WATER, Material.BONE_COMPACT)))
    return SolverConfig(
        geometry=geometry,
        beam=Beam(n_protons=1e6, energy=150.0, energy_spread=1.0, sigma=(0.002, 0.002),
                  center=(0.0, 0.0), angular_spread=0.0, cutoff_sigma=5.0),
        physics=Physics(scattering=Scattering.WENTZEL, nuclear=True),
        numerics=Numerics(spatial_scheme=SpatialScheme.WENO_Z,
                          energy_scheme=EnergyScheme.IF_SSPRK3, wave_speeds=WaveSpeeds.EXACT,
                          max_energy_step=0.25, cfl=0.9, energy_min=1e-4,
                          energy_window_sigma=5.0),
    )
