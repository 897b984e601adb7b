"""Write the golden fixtures under tests/golden/ from the reference's data files.

The reference ships debug dumps of its own GPU pipeline in gpu-tracking/:
  * ``inp``            524x800 float32 LE: the grey_lion.tiff frame (u8 values)
  * ``masses.bin``     524x800 float32 LE: the GPU ``masses`` buffer for that
                       frame (diameter 9, smoothing_size 9, noise_size 1)
  * ``mine_processed`` 2x512x512 float32 LE: plane 0 is a GPU ``processed``
                       frame (plane 1 is zeros)
  * ``centers.bin``    512x512x2 float32 LE: GPU centroid offsets (row, col)
                       computed from that processed frame

Only data is taken (read as raw bytes); no reference code is imported or run.
Run in a container where the reference is mounted:

    python tests/golden/make_fixtures.py /root/reference
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# masses.bin crops: top-left border, bottom-right border, interior block
MASS_CROPS = {
    "tl": (0, 48, 0, 64),
    "br": (476, 524, 736, 800),
    "mid": (200, 264, 300, 364),
}
# mine_processed crop (with a 4-px margin for the disk) and centers crop
PROC_CROP = (100, 236, 100, 236)
CENT_CROP = (104, 232, 104, 232)


def main(ref_root: str) -> None:
    g = os.path.join(ref_root, "gpu-tracking")
    files = ["inp", "masses.bin", "mine_processed", "centers.bin"]
    sha = {f: hashlib.sha256(open(os.path.join(g, f), "rb").read()).hexdigest() for f in files}

    inp = np.fromfile(os.path.join(g, "inp"), dtype="<f4").reshape(524, 800)
    assert np.all(inp == np.round(inp)) and inp.min() >= 0 and inp.max() <= 255
    np.save(os.path.join(HERE, "grey_lion_u8.npy"), inp.astype(np.uint8))

    masses = np.fromfile(os.path.join(g, "masses.bin"), dtype="<f4").reshape(524, 800)
    nz = np.nonzero(masses)
    crops = {}
    for name, (r0, r1, c0, c1) in MASS_CROPS.items():
        crops[name] = masses[r0:r1, c0:c1].copy()
        crops[name + "_box"] = np.array([r0, r1, c0, c1], np.int32)
    crops["nonzero_bbox"] = np.array([nz[0].min(), nz[0].max(), nz[1].min(), nz[1].max()], np.int32)
    crops["nonzero_count"] = np.array([len(nz[0])], np.int64)
    np.savez_compressed(os.path.join(HERE, "grey_lion_masses_crops.npz"), **crops)

    proc = np.fromfile(os.path.join(g, "mine_processed"), dtype="<f4").reshape(2, 512, 512)[0]
    cent = np.fromfile(os.path.join(g, "centers.bin"), dtype="<f4").reshape(512, 512, 2)
    r0, r1, c0, c1 = PROC_CROP
    q0, q1, d0, d1 = CENT_CROP
    np.savez_compressed(
        os.path.join(HERE, "processed_centers_crop.npz"),
        processed=proc[r0:r1, c0:c1].copy(),
        centers=cent[q0:q1, d0:d1].copy(),
        proc_box=np.array(PROC_CROP, np.int32),
        cent_box=np.array(CENT_CROP, np.int32),
    )
    manifest = {
        "source": "amfaber/tracking_optimizations gpu-tracking/ data files",
        "sha256": sha,
        "grey_lion_u8.npy": "inp as uint8 (values are integers 0..255)",
        "grey_lion_masses_crops.npz": {"crops": MASS_CROPS, "params": {"diameter": 9, "smoothing_size": 9, "noise_size": 1}},
        "processed_centers_crop.npz": {"proc_box": PROC_CROP, "cent_box": CENT_CROP, "diameter": 9},
    }
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
