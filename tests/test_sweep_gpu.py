"""Radiance-cache hyper-parameter sweep (experiments.grid_search, restating
gridsearch.h:39-314): every configuration is scored by its mean training
loss over the training-only frames from `score_from` on; results come back
sorted, best first, with every configuration trained for the same frames.
Scores are not bit-reproducible between runs: the training records are
appended in a run-dependent order (as the reference's atomic counter appends
them, RenderKernel.cu / render.h:461-464), so the seeded batch indices pick
different records."""
import numpy as np
import pytest

from tfg_amd import experiments, scenes

pytestmark = pytest.mark.gpu


def test_sweep_scores_sorted(torch_cuda):
    torch = torch_cuda
    sc = scenes.cornell(level=2)
    grid = {"lr": (1e-2, 1e-3), "ema_decay": (0.99,), "steps_per_frame": (1, 2)}
    res = experiments.grid_search(torch, sc, 64, 48, 5, 2, grid=grid, frames=12, score_from=2)
    assert len(res) == 4
    assert {(r["lr"], r["steps_per_frame"]) for r in res} == {(1e-2, 1), (1e-2, 2), (1e-3, 1), (1e-3, 2)}
    scores = [r["score"] for r in res]
    assert scores == sorted(scores) and all(np.isfinite(scores)) and scores[0] > 0
    for r in res:
        assert r["steps"] == 12 * r["steps_per_frame"]
        assert r["ema_decay"] == 0.99
