"""Write ``sample3.npz``: the DIC Challenge "Sample3" images that the
reference's own validation tests read (``tests/test_validation_imageserie_*``,
``doc/00_Validation_ImageSerie.py``), as uint8 arrays.

Kept: the reference image and the frames with rigid shifts of 0.1, 0.5 and
1.0 px in x and y (the shift is in the file name).  The reference's ROI file
(``roi.tif``: grey + alpha, alpha 255 everywhere) gives an all-true mask
through its ``_load_mask`` (mean over channels > 0), so the fixture stores no
mask.  Run from the repository root in a tree where the reference is present:

    python tests/golden/make_sample3.py /path/to/reference

Only data files are read; no reference code is imported or run.
"""
from __future__ import annotations

import re
import sys
from pathlib import Path

import numpy as np
from PIL import Image

HERE = Path(__file__).resolve().parent
KEEP = ("Sample3-001", "Sample3-005", "Sample3-010")


def main(ref_root: str) -> None:
    d = Path(ref_root) / "doc" / "img" / "Sample3"
    ref = np.asarray(Image.open(d / "Sample3 Reference.tif"))
    frames, shifts, names = [], [], []
    for p in sorted(d.glob("Sample3-*.tif")):
        if not p.name.startswith(KEEP):
            continue
        m = re.search(r"X(-?\d+\.\d+)\s+Y(-?\d+\.\d+)", p.stem)
        frames.append(np.asarray(Image.open(p)))
        shifts.append((float(m.group(1)), float(m.group(2))))
        names.append(p.stem)
    assert ref.dtype == np.uint8 and ref.shape == (512, 512) and len(frames) == len(KEEP)
    np.savez_compressed(HERE / "sample3.npz", ref=ref, frames=np.stack(frames),
                        shifts=np.asarray(shifts, np.float64), names=np.asarray(names))
    print("wrote", HERE / "sample3.npz")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
