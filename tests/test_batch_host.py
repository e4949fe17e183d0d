"""Host logic of BatchMeshBased.sequence (reference
``d2ic/batch_mesh_based.py:130-200``) on the CPU, with a recording stand-in for
the DIC pipelines: warm starts from the previous frame or a propagator
(``propagator_previous.py``, ``propagator_constant_velocity.py``), copies of
the warm start, local refinement chained on the global result, and the
per-frame NPZ layout.  The stand-in's u after frame k is (initial guess or
zero) + (k + 1) * step, so every warm start has a closed form."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from d2ic_amd import _lib
from d2ic_amd.batch_mesh_based import BatchMeshBased
from d2ic_amd.dataclasses import BatchConfig, DICDiagnostics, DICResult
from d2ic_amd.propagator_constant_velocity import ConstantVelocityPropagator
from d2ic_amd.propagator_previous import PreviousDisplacementPropagator

NN = 5
STEP = torch.tensor([0.25, -0.5])


class Recorder:
    """Duck-typed DICMeshBased: records prepare / set_initial_guess / run."""

    def __init__(self, add=True, with_disc=False):
        self.guesses, self.frames, self.prepared = [], [], 0
        self.u0 = None
        self.add = add
        self.with_disc = with_disc

    def prepare(self, ref, assets):
        self.prepared += 1

    def set_initial_guess(self, u):
        self.u0 = u
        self.guesses.append((len(self.frames), u))  # the object handed over, not a copy

    def run(self, img):
        k = int(img[0, 0])
        self.frames.append(k)
        base = self.u0 if self.u0 is not None else torch.zeros(NN, 2)
        u = base + (k + 1) * STEP if self.add else base * 1.0
        maps = {"discrepancy_ref": torch.full((2, 3), float(k))} if self.with_disc else {}
        return DICResult(u_nodal=u, strain=torch.zeros(NN, 3), diagnostics=DICDiagnostics(), pixel_maps=maps)


@pytest.fixture(autouse=True)
def cpu_device(monkeypatch):
    monkeypatch.setattr(_lib, "device", lambda: torch.device("cpu"))


def _frames(n):
    return [np.full((4, 4), k, np.float32) for k in range(n)]


def test_warm_start_from_previous_and_none():
    g = Recorder()
    out = BatchMeshBased(np.zeros((4, 4)), None, g, BatchConfig(warm_start_from_previous=True)).run(_frames(3))
    assert g.prepared == 1 and g.frames == [0, 1, 2]
    # frame 0 has no guess; frame k starts from frame k-1's result
    assert [k for k, _ in g.guesses] == [1, 2]
    torch.testing.assert_close(g.guesses[0][1], out.results[0].u_nodal)
    torch.testing.assert_close(g.guesses[1][1], out.results[1].u_nodal)
    torch.testing.assert_close(out.results[2].u_nodal, (1 + 2 + 3) * STEP.expand(NN, 2))
    assert out.diagnostics.info["n_frames"] == 3
    g2 = Recorder()
    BatchMeshBased(np.zeros((4, 4)), None, g2, BatchConfig(warm_start_from_previous=False)).run(_frames(3))
    assert g2.guesses == []


def test_propagators():
    g = Recorder()
    out = BatchMeshBased(np.zeros((4, 4)), None, g, BatchConfig(warm_start_from_previous=False),
                         propagator=ConstantVelocityPropagator()).run(_frames(4))
    u = [r.u_nodal for r in out.results]
    assert [k for k, _ in g.guesses] == [1, 2, 3]
    torch.testing.assert_close(g.guesses[0][1], u[0])  # one frame known: u_prev
    torch.testing.assert_close(g.guesses[1][1], 2.0 * u[1] - u[0])
    torch.testing.assert_close(g.guesses[2][1], 2.0 * u[2] - u[1])
    g2 = Recorder()
    out2 = BatchMeshBased(np.zeros((4, 4)), None, g2, BatchConfig(warm_start_from_previous=False),
                          propagator=PreviousDisplacementPropagator()).run(_frames(3))
    torch.testing.assert_close(g2.guesses[1][1], out2.results[1].u_nodal)


def test_warm_start_is_a_copy_and_local_chain(tmp_path):
    g, loc = Recorder(), Recorder(add=False, with_disc=True)
    cfg = BatchConfig(warm_start_from_previous=True, save_per_frame=True, per_frame_dir=str(tmp_path))
    out = BatchMeshBased(np.zeros((4, 4)), None, g, cfg, dic_local=loc).run(_frames(2))
    assert loc.prepared == 1 and loc.frames == [0, 1]
    # the local stage starts from the global result of the same frame and its
    # result is the one kept (and the next frame's warm start)
    assert [k for k, _ in loc.guesses] == [0, 1]
    torch.testing.assert_close(out.results[0].u_nodal, STEP.expand(NN, 2))
    torch.testing.assert_close(g.guesses[0][1], out.results[0].u_nodal)
    # a copy: changing the stored result does not change the handed-over warm start
    assert g.guesses[0][1].data_ptr() != out.results[0].u_nodal.data_ptr()
    out.results[0].u_nodal.add_(100.0)
    assert float(g.guesses[0][1].abs().max()) < 1.0
    for k in range(2):
        z = np.load(tmp_path / f"frame_{k:04d}.npz")
        assert sorted(z.files) == ["discrepancy_ref", "strain", "u_nodal"]
        assert z["strain"].shape == (NN, 3) and np.all(z["discrepancy_ref"] == k)
    with pytest.raises(RuntimeError):
        BatchMeshBased(np.zeros((4, 4)), None, Recorder(), BatchConfig()).sequence(_frames(1))
