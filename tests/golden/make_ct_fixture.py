"""Cut the section of the reference's CT benchmark into a small fixture.

The reference ships a prostate CT (data/ct/prostate.npz: Hounsfield units,
183 x 183 x 90 voxels of 3 mm, CORT dataset, CC0) and the result of its own
FP64 march on a section of it (thesis_results/scattering_nuclear/ct/
ct_metrics.json).  That section is described there and in the benchmark's
docstring: solver axes (x, y, z) are source axes (0, 2, 1); the beam column
is (x, y) = (80, 45); the skin is the first voxel of that column with
HU > -300; the section spans 10 x 10 voxels around the column and 50 voxels
from the skin; voxels before the first body voxel (HU > -300) of their own
column are outside the body; the march runs on the section refined 6x by
voxel repetition (0.5 mm cells).

This script cuts the 10 x 10 x 50 section (HU, body mask) out of the NPZ
(numpy.load with allow_pickle=False), checks it against the facts
ct_metrics.json records (source digest and shape, skin index, outside-body
fraction, material fractions) and writes

    ct_section.npz   HU (int16) and body_mask (bool), 10 x 10 x 50
    ct_golden.json   the section facts and the reference's dose metrics

No reference code is imported or run.  Run it in a checkout that has the
reference at REFERENCE_ROOT (default /root/reference).
"""

import hashlib
import json
import os
from pathlib import Path

import numpy as np

ROOT = Path(os.environ.get("REFERENCE_ROOT", "/root/reference"))
HERE = Path(__file__).resolve().parent
SOURCE = ROOT / "data" / "ct" / "prostate.npz"
METRICS = ROOT / "thesis_results" / "scattering_nuclear" / "ct" / "ct_metrics.json"

SKIN_HU = -300.0
COLUMN = (80, 45)     # (x, y) of the beam in solver axes
HALF = 5              # lateral half-width, voxels
DEPTH = 50            # voxels from the skin
BONE_FROM_HU = 150.0


def cut(volume_source):
    vol = np.transpose(volume_source, (0, 2, 1))        # solver (x, y, z)
    cx, cy = COLUMN
    col = vol[cx, cy, :]
    skin = int(np.flatnonzero(col > SKIN_HU)[0])
    sec = vol[cx - HALF:cx + HALF, cy - HALF:cy + HALF, skin:skin + DEPTH]
    assert sec.shape == (2 * HALF, 2 * HALF, DEPTH)
    body = np.maximum.accumulate(sec > SKIN_HU, axis=2)
    return sec, body, skin


def main():
    digest = hashlib.sha256(SOURCE.read_bytes()).hexdigest()
    data = np.load(SOURCE, allow_pickle=False)   # an .npy array under an .npz name
    if isinstance(data, np.ndarray):
        hu = data
    else:
        with data:
            hu = np.asarray(data["HU" if "HU" in data.files else data.files[0]])
    ref = json.loads(METRICS.read_text())
    sec_info = ref["section"]
    assert digest == sec_info["source_sha256"], "not the reference's CT"
    assert list(hu.shape) == sec_info["source_shape"]
    sec, body, skin = cut(hu)
    assert skin == sec_info["skin_index_on_beam_axis"]
    outside = float(1.0 - body.mean())
    assert abs(outside - sec_info["outside_body_voxels_fraction"]) < 1e-12
    bone = float(((sec >= BONE_FROM_HU) & body).mean())
    assert abs(bone - ref["material_fractions"]["BONE_COMPACT"]) < 1e-12

    np.savez(HERE / "ct_section.npz", HU=sec.astype(np.int16), body_mask=body)
    m = ref["metrics"]
    golden = {
        "source": "thesis_results/scattering_nuclear/ct/ct_metrics.json",
        "section": sec_info,
        "material_fractions": ref["material_fractions"],
        "hu_partition": ref["hu_partition"],
        "maximum_dose_Gy": m["maximum_dose"]["deterministic_Gy"],
        "r80_mm": m["range"]["deterministic_r80_mm"],
        "dfw_mm": m["range"]["deterministic_dfw_mm"],
        "peak_search_from_mm": m["range"]["peak_search_from_mm"],
        "central_axis_peak_depth_mm": m["central_axis_bragg_peak"]["deterministic_depth_mm"],
        "runtime_seconds": m["performance"]["deterministic_runtime_seconds"],
    }
    (HERE / "ct_golden.json").write_text(json.dumps(golden, indent=1, sort_keys=True) + "\n")
    print("section", sec.shape, "skin", skin, "outside", outside, "bone", bone)


if __name__ == "__main__":
    main()
