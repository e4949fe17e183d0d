"""Synthetic workloads of the benchmark configurations, shared by ``bench.py``
and the GPU tests (no counterpart in the reference, whose benchmark script
reads a fixed image series).

* :func:`plate_sequence` -- 2048x2048 speckle (seed 0), plate with hole (1800^2
  minus a disc of radius 300), frames t = 1..n of u(x,y,t) = t [eps (x - x0),
  -0.3 eps (y - y0)] + 1.5 sin(2 pi x / W) t / 20 on u_y, eps = 1e-3, N(0,1)
  noise; Q4-16 mesh (16,641 nodes, 2,923,392 ROI pixels).
* :func:`rect_single` -- 512x512 speckle (seed 0), centred 400x400 ROI, rigid
  u = (0.37, -0.52) px, N(0,1) noise; Q4-16 (1,089 nodes); one frame.
* :func:`batch_problems` -- n frame pairs of uneven size: sides from
  {512, 1024, 2048, 4096} with weights {.4, .3, .2, .1}, elements from
  {8, 16, 32} px, rectangle or hole ROI, random affine + sinusoid (gradient
  <= 5e-3, amplitude <= 3 px), ZNCC (search 4) + IDW initial guess.
* :func:`serve_streams` -- camera streams with Poisson arrivals, ROI side
  log-uniform in [256, 4096], elements 8..32 px, smooth per-frame motion, ZNCC
  initial guess on each stream's first frame.

Images are rendered on the GPU (``speckle.speckle_sequence_device``) and
stay there.
"""
from __future__ import annotations

import numpy as np

PLATE = dict(H=2048, W=2048, side=1800, hole=300.0, e=16, seed=0, max_iters=200, tol=1e-3, reg=0.1,
             interp="cubic")
RECT = dict(H=512, W=512, side=400, e=16, seed=0, ux=0.37, uy=-0.52, win=41, search=4, max_iters=200, tol=1e-3,
            reg=0.1, interp="cubic")


def plate_sequence(n_frames: int):
    """(ref, frames, mask) of the 4 Mpx plate workload, frames t = 1..n_frames (device tensors)."""
    from . import speckle

    w = PLATE
    fns = [speckle.affine_sinusoid(t, w["W"], w["H"]) for t in range(1, n_frames + 1)]
    ref, frames = speckle.speckle_sequence_device(w["H"], w["W"], w["seed"], fns, noise_sigma=1.0)
    mask = speckle.plate_with_hole_mask(w["H"], w["W"], w["side"], w["hole"])
    return ref, frames, mask


def rect_single():
    """(ref, def, mask) of the 512x512 rigid-translation workload (numpy float32; the
    numpy renderer, so the frame is identical on every machine)."""
    from . import speckle

    w = RECT
    ref, de = speckle.speckle_pair(w["H"], w["W"], w["seed"], speckle.rigid(w["ux"], w["uy"]), noise_sigma=1.0)
    mask = speckle.rect_mask(w["H"], w["W"], w["side"], w["side"])
    return ref, de, mask


def zncc_init(nodes_xy, centers_xy, ref, de, search=4, win=41, max_centres=4096):
    """ZNCC shifts at (a regular subsample of) the element centres, IDW to the
    nodes: the initial guess shared by the plain and the batched/served paths."""
    from .dic_init_motion import project_centers_to_nodes_idw
    from .solver_translation_zncc import match_centers

    c = np.asarray(centers_xy, dtype=np.float32)
    step = max(1, int(np.ceil(len(c) / max_centres)))
    c = c[::step]
    u, _ = match_centers(c, ref, de, win=win, search=search)
    return project_centers_to_nodes_idw(nodes_xy, c, u)


def random_motion(rng, H, max_grad=5e-3, max_amp=3.0):
    """Random affine + sinusoid parameters with |grad u| <= max_grad and |u| <= max_amp."""
    A = rng.uniform(-max_grad, max_grad, size=(2, 2)) / 2.0
    half = H * 0.45
    lin = np.abs(A).sum(axis=1).max() * half
    amp = rng.uniform(0.2, 1.0)
    bmax = max(0.0, max_amp - lin - amp)
    b = rng.uniform(-bmax, bmax, size=2) / np.sqrt(2.0)
    wl = float(H) * rng.uniform(0.8, 2.0)
    # keep the sinusoid's gradient within max_grad as well
    amp = min(amp, max_grad * wl / (2 * np.pi))
    return A, b, amp, wl


def problem(H, e, roi, A, b, amp, wl, seed, interp="cubic"):
    """One frame pair: speckle of side H, ROI, Q4 mesh, pixel assets and a
    CGPlan with f sampled; returns a dict (device tensors)."""
    import torch

    from . import speckle
    from .mask2mesh import mask_to_mesh
    from .mesh_assets import compute_element_centers
    from .pixel_assets import build_pixel_assets
    from .solver_global_cg import CGPlan

    W = H
    if roi == "hole":
        mask = speckle.plate_with_hole_mask(H, W, int(0.88 * H), 0.15 * H)
    else:
        mask = speckle.rect_mask(H, W, int(0.8 * H), int(0.8 * H))
    mesh = mask_to_mesh(mask, e)
    fn = speckle.affine_field(A, b, amp, wl, W / 2.0, H / 2.0)
    ref, frames = speckle.speckle_sequence_device(H, W, seed, [fn], noise_sigma=1.0)
    pa = build_pixel_assets(mesh, ref, 1.0, roi_mask=mask)
    plan = CGPlan(pa)
    plan.set_reference(ref, interp)
    u0 = zncc_init(mesh.nodes_xy, compute_element_centers(mesh), ref, frames[0])
    torch.cuda.synchronize()
    return dict(H=H, e=e, roi=roi, plan=plan, img=frames[0], u0=u0, np=plan.n_pixels, nn=plan.n_nodes, ref=ref,
                mesh=mesh, mask=mask, field=fn)


def batch_problems(seed=7, n=64):
    """The uneven 64-problem batch (see the module docstring)."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([512, 1024, 2048, 4096], size=n, p=[0.4, 0.3, 0.2, 0.1])
    elems = rng.choice([8, 16, 32], size=n)
    rois = rng.choice(["rect", "hole"], size=n)
    probs = []
    for i in range(n):
        A, b, amp, wl = random_motion(rng, int(sizes[i]))
        probs.append(problem(int(sizes[i]), int(elems[i]), str(rois[i]), A, b, amp, wl, seed * 1000 + i))
    return probs


def poisson_arrivals(rng, n_streams, rate, seconds):
    """Sorted (time, stream) pairs: one Poisson process per stream, aggregate ``rate`` per second."""
    arrivals = []
    for s_ in range(n_streams):
        t = rng.exponential(n_streams / rate)
        while t < seconds:
            arrivals.append((t, s_))
            t += rng.exponential(n_streams / rate)
    arrivals.sort()
    return arrivals


def serve_streams(seed=11, n_streams=16, rate=200.0, seconds=1.0, frames_per_stream=None):
    """Streams of the serving workload and their arrivals.

    Returns (streams, arrivals): each stream is a dict with its side, element
    size, pixel assets, reference image, frames (one per arrival of that
    stream, or ``frames_per_stream``), mesh and ZNCC+IDW initial guess."""
    import torch

    from . import speckle
    from .mask2mesh import mask_to_mesh
    from .mesh_assets import compute_element_centers
    from .pixel_assets import build_pixel_assets

    rng = np.random.default_rng(seed)
    arrivals = poisson_arrivals(rng, n_streams, rate, seconds)
    nfr = [sum(1 for _, k in arrivals if k == s_) for s_ in range(n_streams)]
    if frames_per_stream is not None:
        nfr = [max(n, int(frames_per_stream)) for n in nfr]
    streams = []
    for s_ in range(n_streams):
        side = int(np.exp(rng.uniform(np.log(256), np.log(4096))))
        side -= side % 2
        H = side + 64
        e = int(rng.integers(8, 33))
        mask = speckle.rect_mask(H, H, side, side)
        mesh = mask_to_mesh(mask, e)
        A, b, amp, wl = random_motion(rng, side, max_grad=5e-3 / max(nfr[s_], 1), max_amp=3.0)
        fns = [speckle.affine_field(A * (f + 1), b * (f + 1) / max(nfr[s_], 1), amp * (f + 1) / max(nfr[s_], 1),
                                    wl, H / 2.0, H / 2.0) for f in range(max(nfr[s_], 1))]
        ref, frames = speckle.speckle_sequence_device(H, H, seed * 100 + s_, fns, noise_sigma=1.0)
        pa = build_pixel_assets(mesh, ref, 1.0, roi_mask=mask)
        u0 = zncc_init(mesh.nodes_xy, compute_element_centers(mesh), ref, frames[0])
        streams.append(dict(side=side, e=e, pa=pa, ref=ref, frames=frames, u0=u0, mesh=mesh, mask=mask))
    torch.cuda.synchronize()
    return streams, arrivals
