"""Write the regression fixtures in this folder from the CPU oracle.

Run from the repository root:  python tests/golden/make_golden.py

Inputs are small synthetic cases (jittered-grid mesh, speckle pairs from the
seeded spot generator); outputs come from the oracle (the CPU restatement of
the reference algorithms), FP32 and FP64 instances.  Nothing here imports or
runs the reference.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from d2ic_amd import speckle  # noqa: E402
from d2ic_amd.mask2mesh import mask_to_mesh  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests._cases import distorted_mesh  # noqa: E402


def assets_case():
    mesh = distorted_mesh()
    a = O.pixel_assets(mesh.nodes_xy, mesh.elements, (80, 80))
    np.savez_compressed(HERE / "assets_distorted.npz", nodes=mesh.nodes_xy, elements=mesh.elements,
                        image_shape=np.array([80, 80]), **{k: v for k, v in a.items() if isinstance(v, np.ndarray)})


def cg_case():
    m = speckle.rect_mask(64, 64, 48, 48)
    mesh = mask_to_mesh(m, 16)
    ref, de = speckle.speckle_pair(64, 64, 11, speckle.rigid(0.41, -0.23), noise_sigma=1.0)
    a = O.pixel_assets(mesh.nodes_xy, mesh.elements, ref.shape, roi_mask=m)
    out = dict(nodes=mesh.nodes_xy, elements=mesh.elements, roi_mask=m.astype(np.uint8), ref=ref, de=de,
               max_iter=np.int32(20), tol=np.float64(1e-9), reg=np.float64(0.1))
    u0 = np.zeros((mesh.nodes_xy.shape[0], 2), np.float32)
    for interp in ("linear", "cubic"):
        for dt, tag in ((np.float32, "f32"), (np.float64, "f64")):
            u, h, n = O.cg_solve(u0, ref, de, a, alpha_reg=0.1, max_iter=20, tol=1e-9, interp=interp, dtype=dt)
            out[f"u_{interp}_{tag}"] = u
            out[f"hist_{interp}_{tag}"] = h
            out[f"n_{interp}_{tag}"] = np.int32(n)
    np.savez_compressed(HERE / "cg_rect64.npz", **out)


def zncc_case():
    ref, de = speckle.speckle_pair(96, 96, 7, speckle.rigid(2.6, -3.3), noise_sigma=1.0)
    rng = np.random.default_rng(2)
    centers = rng.uniform(8, 88, size=(32, 2)).astype(np.float32)
    u, s = O.zncc(centers, ref, de, 15, 4)
    np.savez_compressed(HERE / "zncc_speckle96.npz", ref=ref, de=de, centers=centers, win=np.int32(15),
                        search=np.int32(4), u=u, scores=s)


if __name__ == "__main__":
    O.build()
    assets_case()
    cg_case()
    zncc_case()
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size)
