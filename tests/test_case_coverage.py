"""The GPU parity cases exercise both ways the FP64 kernel handles the mesh
metrics of the identity map: grids whose rounded metrics (kappa, gamma_xi,
gamma_eta of src/domain/SMOD_mesh_manipulation.f90) are all exactly 1,
which the kernel skips, and grids where they are not, which it computes."""
from __future__ import annotations

import numpy as np

from tests.test_gpu_parity import BATCH_SIZES, CASES


def _unit(O, nx, ny):
    m = O.metrics(O.Config(nx=nx, ny=ny))
    return all(bool(np.all(m[k] == 1.0)) for k in ("kappa", "gamma_xi", "gamma_eta"))


def test_single_instance_cases_cover_both_metric_kinds(oracle):
    kinds = {_unit(oracle, c[0], c[1]) for c in CASES}
    assert kinds == {True, False}


def test_batch_mixes_both_metric_kinds(oracle):
    kinds = {_unit(oracle, nx, ny) for nx, ny, _ in BATCH_SIZES}
    assert kinds == {True, False}
