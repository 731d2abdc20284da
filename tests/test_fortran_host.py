"""The Fortran host (paraclaws_amd/fortran): USER_DIR override rules, the
control.in reader, and (on the GPU) a full run against the oracle."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "paraclaws_amd", "fortran")
TOOL = os.path.join(FDIR, "mk", "source_list.py")
EXAMPLE = os.path.join(FDIR, "examples", "dam_break")
EXE = os.path.join(EXAMPLE, "build", "paraclaws_gpu")


def base_sources():
    out = subprocess.run(["make", "-s", "-C", FDIR, "sources"], capture_output=True, text=True,
                         check=True).stdout.split()
    return out


def source_list(user_dir, base):
    return subprocess.run([sys.executable, TOOL, "--base-root", FDIR, "--user-dir", str(user_dir), *base],
                          capture_output=True, text=True)


def test_override_replaces_in_place_and_appends_new(tmp_path):
    base = base_sources()
    (tmp_path / "my_ic.f90").write_text(
        "submodule(MOD_user_subroutines) SMOD_set_initial_conditions\nend submodule\n")
    (tmp_path / "extra.f90").write_text("module my_extra_module\nend module\n")
    (tmp_path / "build").mkdir()
    (tmp_path / "build" / "ignored.f90").write_text("module ignored_mod\nend module\n")
    r = source_list(tmp_path, base)
    assert r.returncode == 0, r.stderr
    out = r.stdout.split()
    k = [i for i, p in enumerate(base) if p.endswith("SMOD_set_initial_conditions.f90")][0]
    assert out[k] == str(tmp_path / "my_ic.f90")
    assert out[-1] == str(tmp_path / "extra.f90")
    assert len(out) == len(base) + 1
    assert not any("ignored" in p for p in out)


def test_override_errors(tmp_path):
    base = base_sources()
    a = tmp_path / "a"
    a.mkdir()
    (a / "one.f90").write_text("submodule(MOD_user_subroutines) SMOD_setup\nend submodule\n")
    (a / "two.f90").write_text("submodule (MOD_user_subroutines)   SMOD_setup ! differs\nend submodule\n")
    assert source_list(a, base).returncode != 0          # two files replace one base file
    b = tmp_path / "b"
    b.mkdir()
    (b / "both.f90").write_text("module MOD_constants\nend module\nmodule MOD_time_control\nend module\n")
    assert source_list(b, base).returncode != 0          # one file replaces two base files
    c = tmp_path / "c"
    c.mkdir()
    txt = "submodule(MOD_user_subroutines) SMOD_setup\nend submodule\n"
    (c / "x.f90").write_text(txt)
    (c / "y.f90").write_text(txt)
    assert source_list(c, base).returncode == 0          # identical duplicates are tolerated


def test_host_builds_with_example_override():
    r = subprocess.run(["make", "-s", "-C", FDIR, f"USER_DIR={EXAMPLE}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(EXE)


def _control_in(n, t_final, limiter=4, bc=(4, 4, 4, 4), cfl=0.9, dt0=1e-10, ho=True, lim=True):
    vals = [3, 2, 2, 10, 2, 0, t_final, 10, dt0, ".true.", cfl, 1e-10, 1e-2, 1.2, 0.5,
            ".true." if ho else ".false.", ".true." if lim else ".false.", limiter, n, n, -1, 1, -1, 1,
            *bc]
    return "\n".join(str(v) for v in vals) + "\n"


@pytest.mark.gpu
@pytest.mark.parametrize("precision,engine", [(0, 0), (0, 1)])
def test_fortran_host_run_matches_oracle(oracle, tmp_path, precision, engine):
    """The Fortran program (control.in, custom.in, dam-break override) on the
    GPU reproduces the oracle's run bit for bit (FP64)."""
    if not os.path.exists(EXE):
        pytest.skip("Fortran host not built (run __graft_entry__.build())")
    n = 64
    (tmp_path / "control.in").write_text(_control_in(n, 0.05))
    (tmp_path / "custom.in").write_text(f"{precision}\n{engine}\n")
    print(f"fortran host run precision={precision} engine={engine}", flush=True)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    got = np.fromfile(tmp_path / "output" / "final_state.bin", dtype=np.float64)
    got = got.reshape(n + 4, n + 4, 3)
    cfg = oracle.Config(nx=n, ny=n, limiter=4, bc=(4, 4, 4, 4))
    ts = oracle.TimeSettings(t_final=0.05, cfl_target=0.9, dt_max=1e-2, n_snap=10)
    q = oracle.dam_break(cfg)
    recs, tc = oracle.run(cfg, ts, q)
    steps = [ln for ln in r.stdout.splitlines() if ln.startswith(" paraCLAWs - time:") or ln.startswith("paraCLAWs - time:")]
    assert len(steps) == sum(x[0] for x in recs) + 1   # + the initial line
    from tests.helpers import sides_mask

    m = sides_mask(n, n)
    assert np.array_equal(got[m], q[m])
    snaps = sorted(os.listdir(tmp_path / "output"))
    assert "domain.dat" in snaps and "mesh_rank_0.dat" in snaps
    times = sorted((s for s in snaps if s[0].isdigit()), key=float)
    assert len(times) == 11   # t = 0 and 10 snapshots
    # the last snapshot is the final state, written in the reference's
    # qout.dat format: 3 values per line (50e26.16), j outer, i inner, ghosts included
    qout = np.loadtxt(tmp_path / "output" / times[-1] / "rank_0" / "qout.dat")
    assert qout.shape == ((n + 4) * (n + 4), 3)
    assert np.allclose(qout.reshape(n + 4, n + 4, 3)[m], q[m], rtol=1e-15, atol=1e-16)


@pytest.mark.gpu
def test_fortran_host_stretched_grid_matches_oracle(oracle, tmp_path):
    """The reference's stretched example (x = 4 xi, y = 2 eta, through a
    USER_DIR override of SMOD_coordinate_mapping): the host uploads the
    capacity and length ratios and the run matches the oracle given the same
    metric arrays, bit for bit."""
    from paraclaws_amd import ic

    exe = os.path.join(FDIR, "examples", "dam_break_stretched", "build", "paraclaws_gpu")
    if not os.path.exists(exe):
        pytest.skip("stretched example not built (run __graft_entry__.build())")
    n = 50
    (tmp_path / "control.in").write_text(_control_in(n, 0.05, bc=(2, 2, 2, 2)))
    (tmp_path / "custom.in").write_text("0\n0\n")
    print("fortran host stretched grid", flush=True)
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "mapped grid" in r.stdout
    got = np.fromfile(tmp_path / "output" / "final_state.bin", dtype=np.float64).reshape(n + 4, n + 4, 3)
    k, gx, ge, x, y = ic.mapped_metrics(n, n, lambda a, b: (4 * a, 2 * b))
    cfg = oracle.Config(nx=n, ny=n, limiter=4, bc=(2, 2, 2, 2), kappa=np.ascontiguousarray(k),
                        gamma_xi=np.ascontiguousarray(gx), gamma_eta=np.ascontiguousarray(ge))
    q = cfg.frame()
    q[..., 0] = np.where(np.sqrt(x ** 2 + y ** 2) < 0.5, 2.0, 1.0)
    ts = oracle.TimeSettings(t_final=0.05, cfl_target=0.9, dt_max=1e-2, n_snap=10)
    oracle.run(cfg, ts, q)
    from tests.helpers import sides_mask

    m = sides_mask(n, n)
    assert np.array_equal(got[m], q[m])
