"""Result persistence (persistence.py:13-369), CPU only.

Pinned by the reference's saved runs: tests/golden/thesis_golden.json holds
the config.json and metadata config_hash of 30 thesis runs.
  * the canonical hash of every stored config.json is its recorded hash;
  * the configuration rebuilt from a stored config.json serialises to the
    same dictionary, up to the fields the reference's dataclasses gained or
    lost since those runs (listed below), and therefore to the same hash;
  * a saved run has the reference's directory name, files and keys, and its
    arrays load without pickling.
"""

import json

import numpy as np
import pytest

from helpers import ct_config, ct_section_npz, thesis_cases
from proton_solver import (Backend, Beam, Box, CoefficientInterpolation, EnergyScheme, Geometry,
                           Material, Numerics, OutputConfig, Physics, Precision, Reference,
                           References, Scattering, SolverConfig, SpatialScheme, WaveSpeeds)
from proton_solver import persistence as P
from proton_solver.config import DeviceOptions
from proton_solver.enums import Backend as B
from proton_solver.numerics.grid import Grid
from proton_solver.result import Result

# Fields of the stored configs that the current dataclasses no longer have,
# and fields added since (with their serialised defaults).
REMOVED = {("numerics", "active_domain"), ("numerics", "active_relative_tolerance")}
ADDED = {("physics", "coefficient_interpolation"), ("geometry", "ct_scan"),
         ("geometry", "hu_partition")}


def _config_from_json(c):
    g, b, p, n, r = c["geometry"], c["beam"], c["physics"], c["numerics"], c["references"]
    objects = [Box(o["x_min"], o["x_max"], o["y_min"], o["y_max"], o["z_min"], o["z_max"],
                   Material[o["material"]]) for o in g["objects"]]
    num = {k: v for k, v in n.items() if ("numerics", k) not in REMOVED}
    for k in ("spatial_scheme", "energy_scheme", "wave_speeds"):
        num[k] = {"spatial_scheme": SpatialScheme, "energy_scheme": EnergyScheme,
                  "wave_speeds": WaveSpeeds}[k][num[k]]
    ref = dict(r)
    ref["reference"] = Reference[ref["reference"]]
    ref["additional_references"] = tuple(Reference[x] for x in ref["additional_references"])
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
        numerics=Numerics(**num),
        backend=Backend[c["backend"]],
        solve_deterministic=c["solve_deterministic"],
        references=References(**ref),
    )


CASES = thesis_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["source"].split("thesis_results/")[1] for c in CASES])
def test_stored_configs_hash_and_round_trip(case):
    stored = case["config_json"]
    assert P.hash_config_dict(stored) == case["config_hash"]
    ours = P.config_dict(_config_from_json(stored))
    for sec, key in ADDED:
        if key not in stored[sec]:
            ours[sec].pop(key)
    for sec, key in REMOVED:
        if key in stored[sec]:
            ours[sec][key] = stored[sec][key]
    assert ours == stored
    assert P.hash_config_dict(ours) == case["config_hash"]


def test_added_fields_serialise_like_the_reference(tmp_path):
    c = SolverConfig(geometry=Geometry(size=(0.01, 0.01, 0.02), resolution=(4, 4, 8)),
                     beam=Beam(energy=50.0))
    d = P.config_dict(c)
    assert d["physics"]["coefficient_interpolation"] == "LINEAR"
    assert d["geometry"]["ct_scan"] is None
    assert d["geometry"]["hu_partition"] == {"cutoffs": [-200.0, 150.0],
                                             "materials": ["LUNG", "WATER", "BONE_COMPACT"]}
    assert "output" not in d and "view_geometry" not in d["geometry"] and "device" not in d
    # the device index does not enter the hash; a precision other than FP64 does
    h = P.config_hash(c)
    c1 = SolverConfig(geometry=c.geometry, beam=c.beam, device=DeviceOptions(device=1),
                      output=OutputConfig(results_directory="elsewhere"))
    assert P.config_hash(c1) == h
    c2 = SolverConfig(geometry=c.geometry, beam=c.beam,
                      device=DeviceOptions(precision=Precision.FP32_BULK))
    assert P.config_hash(c2) != h
    assert P.config_dict(c2)["device"]["precision"] == "FP32_BULK"
    # CT volumes enter by digest
    ct = P.config_dict(ct_config(ct_section_npz(tmp_path, upsample=1), upsample=1))
    hu = ct["geometry"]["ct_scan"]["hu"]
    assert hu["shape"] == [10, 10, 50] and hu["dtype"] == "int16" and len(hu["sha256"]) == 64


def test_save_result_layout(tmp_path):
    c = SolverConfig(geometry=Geometry(size=(0.004, 0.005, 0.006), resolution=(4, 5, 6)),
                     beam=Beam(energy=50.0, sigma=(0.001, 0.001)),
                     output=OutputConfig(save_results=True, results_directory=str(tmp_path)))
    grid = Grid.from_geometry(c.geometry, np)
    rng = np.random.default_rng(0)
    res = Result(phi=rng.random(grid.shape), J=rng.random((3, *grid.shape)),
                 dose=rng.random(grid.shape), grid=grid, energies=np.linspace(55, 1, 9),
                 runtime=2.0, bfp_mc_dose=rng.random(grid.shape), bfp_mc_runtime=10.0,
                 backend=B.GPU)
    run = P.save_result(res, c, git_cwd=tmp_path)          # not a git checkout
    parts = run.name.split("__")
    assert parts[1:] == ["4x5x6", "weno_z", P.config_hash(c)[:8], "nogit"]
    assert len(parts[0]) == 15
    cfg = json.loads((run / "config.json").read_text())
    assert cfg == P.config_dict(c)
    meta = json.loads((run / "metadata.json").read_text())
    assert set(meta) == {"created_at", "config_hash", "git_commit", "git_dirty",
                         "git_state_hash", "runtime_seconds", "bfp_mc_runtime_seconds",
                         "topas_runtime_seconds", "speedup_vs_bfp_mc_x", "speedup_vs_topas_x",
                         "backend"}
    assert meta["config_hash"] == P.config_hash(c) and meta["git_commit"] is None
    assert meta["speedup_vs_bfp_mc_x"] == 5.0 and meta["speedup_vs_topas_x"] == "-"
    assert meta["backend"] == "GPU"
    with np.load(run / "results.npz", allow_pickle=False) as d:
        assert sorted(d.files) == sorted(["x", "y", "z", "phi", "J", "dose", "energies",
                                          "bfp_mc_dose"])
        assert np.array_equal(d["dose"], res.dose) and np.array_equal(d["x"], grid.x)


def test_git_provenance_of_a_checkout(tmp_path):
    import subprocess

    def git(*a):
        subprocess.run(["git", *a], cwd=tmp_path, check=True, capture_output=True)

    git("init", "-q")
    (tmp_path / "a.txt").write_text("one")
    git("add", "a.txt")
    git("-c", "user.name=t", "-c", "user.email=t@t", "commit", "-q", "-m", "a")
    clean = P.git_provenance(tmp_path)
    assert clean["git_dirty"] is False and len(clean["git_commit"]) == 40
    (tmp_path / "b.txt").write_text("two")             # untracked file: dirty
    dirty = P.git_provenance(tmp_path)
    assert dirty["git_dirty"] is True
    assert dirty["git_state_hash"] != clean["git_state_hash"]
