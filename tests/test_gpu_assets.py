"""Device pixel assets vs the CPU oracle: integer outputs bit-exact, float outputs bit-exact
on dyadic structured meshes and within 2 ulp elsewhere."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests._cases import distorted_mesh, structured_case

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _compare(pa, ref, exact_float=True):
    for key in ("pixel_nodes", "node_neighbor_index", "node_neighbor_degree", "node_pixel_index",
                "node_pixel_degree", "roi_mask_flat"):
        a = _np(getattr(pa, key))
        b = np.asarray(ref[key])
        assert a.shape == b.shape, key
        assert np.array_equal(a, b.astype(a.dtype)), key
    assert np.array_equal(_np(pa.pixel_elements), ref["pixel_elts"])
    for key in ("pixel_coords_ref", "pixel_shapeN", "node_neighbor_weight", "node_N_weight", "node_reg_weight"):
        a = _np(getattr(pa, key))
        b = np.asarray(ref[key], dtype=np.float32)
        assert a.shape == b.shape, key
        if exact_float:
            assert np.array_equal(a, b), key
        else:
            np.testing.assert_allclose(a, b, rtol=0, atol=4e-7, err_msg=key)


@pytest.mark.parametrize("e,use_roi,hole", [(16, False, 0.0), (16, True, 0.0), (8, True, 20.0), (32, False, 0.0)])
def test_structured_assets_bit_exact(gpu, e, use_roi, hole):
    from d2ic_amd.pixel_assets import build_pixel_assets

    print(f"[case] structured e={e} roi={use_roi} hole={hole}", flush=True)
    m, mesh, ref, _ = structured_case(H=160, W=160, side=128, e=e, hole=hole)
    roi = m if use_roi else None
    pa = build_pixel_assets(mesh, ref, 1.0, roi_mask=roi)
    oa = O.pixel_assets(mesh.nodes_xy, mesh.elements, ref.shape, roi_mask=roi)
    _compare(pa, oa, exact_float=True)


def test_distorted_mesh_assets(gpu):
    from d2ic_amd.pixel_assets import build_pixel_assets

    print("[case] distorted mesh", flush=True)
    mesh = distorted_mesh()
    ref = np.zeros((80, 90), np.float32)
    pa = build_pixel_assets(mesh, ref, 1.0)
    oa = O.pixel_assets(mesh.nodes_xy, mesh.elements, ref.shape)
    _compare(pa, oa, exact_float=False)


def test_reference_fixture_unit_quad(gpu):
    """tests/test_pixel_assets.py:6-36 of the reference, through the device path."""
    import torch

    from d2ic_amd import _lib
    import ctypes

    nodes = torch.tensor([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]], device="cuda")
    el = torch.tensor([[0, 1, 2, 3]], dtype=torch.int32, device="cuda")
    # candidate rectangle rows [0,2) cols [0,2): centres (0.5,0.5) (1.5,0.5) (0.5,1.5) (1.5,1.5)
    cand = torch.empty(4, dtype=torch.int32, device="cuda")
    nk = ctypes.c_int64(0)
    _lib.call("d2ic_pixel_map", _lib.ptr(nodes), 4, _lib.ptr(el), 1, 2, 2, 0, 2, 0, 2, None, _lib.ptr(cand),
              ctypes.byref(nk), _lib.stream())
    assert cand.cpu().tolist() == [0, -1, -1, -1]
    assert nk.value == 1
