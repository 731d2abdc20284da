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


def read_vts(path):
    """Point arrays and extent of a VTK StructuredGrid file with inline
    "binary" DataArrays (base64 UInt32 byte count, then base64 data)."""
    import base64
    import re

    txt = open(path).read()
    out = {"extent": [int(v) for v in re.search(r'<Piece Extent="([^"]+)"', txt).group(1).split()]}
    for name, body in re.findall(r'<DataArray[^>]*Name="([^"]+)"[^>]*format="binary">\s*([A-Za-z0-9+/=]+)\s*</DataArray>', txt):
        nbytes = int(np.frombuffer(base64.b64decode(body[:8]), dtype="<u4")[0])
        data = base64.b64decode(body[8:])
        assert len(data) == nbytes
        out[name] = np.frombuffer(data, dtype="<f8")
    return out


def test_vtk_base64_writer(tmp_path):
    """MOD_IO's base64 writer against Python's, for every length mod 3, and
    a DataArray body (UInt32 byte count encoded apart from the data)."""
    import base64

    build = os.path.join(FDIR, "build")
    if not os.path.exists(os.path.join(build, "mod", "mod_io.mod")):
        pytest.skip("Fortran host not built (run __graft_entry__.build())")
    (tmp_path / "t.f90").write_text("""program t
    use MOD_IO
    use iso_fortran_env, only: int8
    integer(int8) :: b(7) = [0_int8, -1_int8, 127_int8, -128_int8, 64_int8, 3_int8, -77_int8]
    double precision :: a(3) = [1.5d0, -2.25d0, 1d300]
    integer :: u, n
    open(newunit=u, file='out.txt', access='stream', form='unformatted', status='replace')
    do n = 0, 7
        call write_b64(u, b(1:n))
        write(u) new_line('a')
    end do
    call write_b64_array(u, a)
    close(u)
end program
""")
    objs = [os.path.join(build, "obj", f) for f in sorted(os.listdir(os.path.join(build, "obj")))
            if f.endswith(".o") and f != "main.o"]
    lib = os.path.join(ROOT, "paraclaws_amd")
    fc = "/opt/rocm/lib/llvm/bin/flang"
    r = subprocess.run([fc, "-cpp", "-fopenmp", "-module-dir", str(tmp_path), "-I", os.path.join(build, "mod"),
                        "t.f90", *objs, f"-L{lib}", "-lpcw_hip", f"-Wl,-rpath,{lib}", "-o", "t"],
                       cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run(["./t"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lines = (tmp_path / "out.txt").read_text().split("\n")
    raw = bytes([0, 255, 127, 128, 64, 3, 256 - 77])
    for n in range(8):
        assert lines[n] == base64.b64encode(raw[:n]).decode()
    a = np.array([1.5, -2.25, 1e300], dtype="<f8").tobytes()
    assert lines[8] == (base64.b64encode(np.uint32(len(a)).tobytes()) + base64.b64encode(a)).decode()


_TEXTIO_DRIVER = """
program check_textio
    use MOD_textio, only: write_e26_records
    implicit none
    integer, parameter :: ni = 37, nj = 150
    double precision, allocatable :: a(:, :, :)
    integer :: m, i, j, k, u
    character(len=16) :: name
    do m = 3, 10, 7
        allocate(a(m, ni, nj))
        call random_seed(put=[(977 + k, k = 1, 64)])
        call random_number(a)
        a = (a - 0.5d0) * 10.d0 ** (floor(a * 60) - 30)
        a(1, 1, 1) = 0.d0
        a(2, 1, 1) = -0.d0
        a(3, 1, 1) = 4.9406564584124654d-324
        a(1, 2, 1) = 1.7976931348623157d308
        a(2, 2, 1) = 0.99999999999999995d0
        a(3, 2, 1) = 1.d100
        a(1, 3, 1) = -9.81000041961669921875d0
        a(2, 3, 1) = transfer(int(z'7FF8000000000000', 8), 1.d0)
        a(3, 3, 1) = -transfer(int(z'7FF0000000000000', 8), 1.d0)
        write(name, '(a,i0)') 'serial_', m
        open(newunit=u, file=trim(name), status='replace', form='formatted')
        do j = 1, nj
            do i = 1, ni
                write(u, '(50e26.16)') a(:, i, j)
            end do
        end do
        close(u)
        write(name, '(a,i0)') 'parallel_', m
        call write_e26_records(trim(name), a)
        deallocate(a)
    end do
end program
"""


def test_snapshot_text_writer_equals_the_serial_write(tmp_path):
    """MOD_textio.write_e26_records (qout.dat/aout.dat of every snapshot:
    internal writes of (50e26.16) on OpenMP threads, blocks of rows streamed
    to the file) writes exactly the bytes of the reference's serial
    write(u, '(50e26.16)') a(:, i, j) loop: records of 3 and 10 fields, 150
    rows (not a multiple of the 64-row blocks), values over 60 decades with
    signed zeros, a denormal, the largest double, 3-digit exponents, NaN and
    -Inf, on 4 threads."""
    fc = "/opt/rocm/lib/llvm/bin/flang"
    if not os.path.exists(fc):
        pytest.skip("flang not found")
    (tmp_path / "check.f90").write_text(_TEXTIO_DRIVER)
    src = os.path.join(FDIR, "src", "utility", "MOD_textio.f90")
    subprocess.run([fc, "-O2", "-fopenmp", "-o", "check", src, "check.f90"], cwd=tmp_path, check=True,
                   capture_output=True)
    env = dict(os.environ, OMP_NUM_THREADS="4")
    subprocess.run(["./check"], cwd=tmp_path, check=True, env=env)
    for m in (3, 10):
        serial = (tmp_path / f"serial_{m}").read_bytes()
        assert serial.count(b"\n") == 37 * 150
        assert (tmp_path / f"parallel_{m}").read_bytes() == serial


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
    # VTK snapshots in the reference's layout: the piece holds the interior
    # cell centres and q1..q3 exactly; the parallel header names the piece;
    # the collection lists every snapshot
    piece = read_vts(tmp_path / "output" / times[-1] / "rank_0" / "solution_rank_0.vts")
    assert piece["extent"] == [1, n, 1, n, 1, 1]
    xc = -1.0 + (np.arange(n) + 0.5) * (2.0 / n)
    pts = piece["Points"].reshape(n, n, 3)
    assert np.allclose(pts[..., 0], xc[None, :], atol=1e-15) and np.allclose(pts[..., 1], xc[:, None], atol=1e-15)
    for k in range(3):
        assert np.array_equal(piece[f"q{k + 1}"].reshape(n, n), q[2:-2, 2:-2, k])
    pvts = (tmp_path / "output" / times[-1] / "solution.pvts").read_text()
    assert 'Source="rank_0/solution_rank_0.vts"' in pvts and 'Name="q3"' in pvts
    pvd = (tmp_path / "output" / "solution.pvd").read_text()
    assert pvd.count("<DataSet ") == 11 and f'file="{times[-1]}/solution.pvts"' in pvd


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


BATHY_EXAMPLE = os.path.join(FDIR, "examples", "bathymetry")
HOOKS_CASE = os.path.join(ROOT, "tests", "fortran_hooks")
SOURCE_CASE = os.path.join(ROOT, "tests", "fortran_source")


def _exe(user_dir):
    exe = os.path.join(user_dir, "build", "paraclaws_gpu")
    if not os.path.exists(exe):
        pytest.skip(f"Fortran host of {user_dir} not built (run __graft_entry__.build())")
    return exe


def test_device_physics_override_replaces_the_base_submodules():
    """The GPU path's settings are a hook submodule like the reference's: a
    USER_DIR file declaring SMOD_device_physics (or SMOD_setup, SMOD_user_post)
    takes the base file's place in the compile order."""
    base = base_sources()
    for user_dir, names in ((HOOKS_CASE, ["SMOD_setup", "SMOD_user_post", "SMOD_device_physics",
                                          "SMOD_set_initial_conditions"]),
                            (BATHY_EXAMPLE, ["SMOD_device_physics", "SMOD_set_initial_conditions"])):
        r = source_list(user_dir, base)
        assert r.returncode == 0, r.stderr
        got = r.stdout.split()
        assert len(got) == len(base)
        for i, (b, g) in enumerate(zip(base, got)):
            stem = os.path.basename(b)[:-4]
            if stem in names:
                assert g == os.path.join(user_dir, stem + ".f90"), (i, g)
            else:
                assert g == b


@pytest.mark.gpu
@pytest.mark.parametrize("engine,loop", [(0, 0), (1, 0), (0, 1)])
def test_fortran_host_bathymetry_example_matches_oracle(oracle, tmp_path, engine, loop):
    """Config 3's physics selected through the host: the example's
    SMOD_device_physics override (bathymetry slope, wet/dry cut-off) and
    initial state (still water over the seeded Gaussian bathymetry, dam-break
    step on x < 0, dry cells), outflow sides, run by the Fortran program on
    the GPU in FP64, equal the oracle bit for bit, with the time loop on the
    host (both engines) and on the device (custom.in line 3)."""
    from paraclaws_amd import ic

    exe = _exe(BATHY_EXAMPLE)
    nx = ny = 96
    (tmp_path / "control.in").write_text(_control_in(nx, 0.03, bc=(2, 2, 2, 2)))
    (tmp_path / "custom.in").write_text(f"0\n{engine}\n{loop}\n")
    b = ic.gaussian_bathymetry(nx, ny, seed=3)
    b.tofile(tmp_path / "bathymetry.bin")
    print(f"fortran host bathymetry example engine={engine} loop={loop}", flush=True)
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "device source kind 1" in r.stdout
    got = np.fromfile(tmp_path / "output" / "final_state.bin", dtype=np.float64).reshape(ny + 4, nx + 4, 3)
    cfg = oracle.Config(nx=nx, ny=ny, limiter=4, bc=(2, 2, 2, 2), source=1, dry_tol=1e-6)
    ts = oracle.TimeSettings(t_final=0.03, cfl_target=0.9, dt_max=1e-2, n_snap=10)
    q = ic.still_water_with_step(b, nx, dry_tol=1e-6)
    assert (q[..., 0] == 0.0).any()  # the start has dry cells
    recs, tc = oracle.run(cfg, ts, q, bathy=np.ascontiguousarray(b))
    from tests.helpers import sides_mask

    m = sides_mask(nx, ny)
    assert tc.time >= 0.03 - 1e-15
    assert np.array_equal(got[m].view(np.uint64), q[m].view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("loop", [0, 1])
def test_fortran_host_hooks_change_the_device_state(oracle, tmp_path, loop):
    """user_post (through execute_post at start-up and at every snapshot) and
    after_step work on the current state and their changes reach the device:
    the test case's user_post raises h by 1e-3 in every interior cell, its
    after_step zeroes hu after the fifth accepted step.  The oracle replays
    the same run step by step with the same changes between steps; the final
    states are equal bit for bit (FP64).  Asked for the device time loop
    (custom.in line 3 = 1), a case whose after_step needs the state every
    step says so and runs the host loop."""
    exe = _exe(HOOKS_CASE)
    n, t_final = 64, 0.05
    (tmp_path / "control.in").write_text(_control_in(n, t_final))
    (tmp_path / "custom.in").write_text(f"0\n0\n{loop}\n")
    print(f"fortran host hooks loop={loop}", flush=True)
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert ("the loop runs on the host" in r.stdout) == (loop == 1)
    got = np.fromfile(tmp_path / "output" / "final_state.bin", dtype=np.float64).reshape(n + 4, n + 4, 3)
    cfg = oracle.Config(nx=n, ny=n, limiter=4, bc=(4, 4, 4, 4))
    ts = oracle.TimeSettings(t_final=t_final, cfl_target=0.9, dt_max=1e-2, n_snap=10)
    q = oracle.dam_break(cfg)
    oracle.fill_ghosts(cfg, q)         # start-up: update_rank_boundaries
    q[2:-2, 2:-2, 0] += 1e-3           # execute_post at t = 0
    oracle.fill_ghosts(cfg, q)         # the boundary update before the loop
    tc, accepted, posts = None, 0, 1
    while tc is None or tc.time < tc.time_final:
        recs, tc = oracle.run(cfg, ts, q, max_accepted=1, tc=tc if tc is not None else ts.tc())
        assert recs and recs[-1][0] == 1
        accepted += 1
        if recs[-1][1] > 0:            # a snapshot: execute_post
            q[2:-2, 2:-2, 0] += 1e-3
            posts += 1
        if accepted == 5:              # after_step
            q[..., 1] = 0.0
    assert posts == 11
    from tests.helpers import sides_mask

    m = sides_mask(n, n)
    assert np.array_equal(got[m].view(np.uint64), q[m].view(np.uint64))


@pytest.mark.gpu
def test_fortran_host_rejects_a_nonzero_source_function(tmp_path):
    """A user source_function that is nonzero on the initial state, with no
    device source selected, stops the program at start-up with a message:
    the GPU step cannot call it per cell."""
    exe = _exe(SOURCE_CASE)
    (tmp_path / "control.in").write_text(_control_in(32, 0.01))
    (tmp_path / "custom.in").write_text("0\n0\n")
    print("fortran host nonzero source_function", flush=True)
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "source_function returns a nonzero source" in r.stdout + r.stderr
    assert not (tmp_path / "output" / "final_state.bin").exists()


def _run_host(exe, case_dir, control, custom):
    case_dir.mkdir()
    (case_dir / "control.in").write_text(control)
    (case_dir / "custom.in").write_text(custom)
    r = subprocess.run([exe], cwd=case_dir, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    # the tile shape the program picks (timed, so it may differ between runs)
    # does not change the run
    lines = [ln for ln in r.stdout.splitlines() if not ln.startswith(" device tiles:")]
    first = next(k for k, ln in enumerate(lines) if "paraCLAWs - time:" in ln)
    last = next(k for k, ln in enumerate(lines) if ln.startswith("accepted steps:"))
    return lines[first:last], lines[last]


@pytest.mark.gpu
@pytest.mark.parametrize("precision,engine,dt0", [(0, 0, 1e-10), (0, 0, 0.2), (1, 0, 1e-10), (0, 1, 0.2)])
def test_fortran_host_device_loop_equals_host_loop(oracle, tmp_path, precision, engine, dt0):
    """custom.in line 3 = 1 runs the time loop on the device (graph-launched
    attempts under the device's copy of the controller, back on the host at
    every snapshot, each run's attempts replayed through the host's
    time_control).  Against the host loop (line 3 = 0) on the same case: the
    same terminal output line for line (step lines, rejection messages and
    snapshot lines), the same accepted and rejected counts, the same snapshot
    directories and the same final state bit for bit, for the fused engines
    and the staged one (engine 1); dt0 = 0.2 starts with a rejected attempt
    (CFL far above the target).  FP64 with dt0 = 1e-10 is also the oracle's
    run bit for bit."""
    if not os.path.exists(EXE):
        pytest.skip("Fortran host not built (run __graft_entry__.build())")
    n, t_final = 64, 0.05
    control = _control_in(n, t_final, dt0=dt0)
    print(f"fortran host loops precision={precision} engine={engine} dt0={dt0}", flush=True)
    out_h, tail_h = _run_host(EXE, tmp_path / "host", control, f"{precision}\n{engine}\n0\n")
    out_d, tail_d = _run_host(EXE, tmp_path / "device", control, f"{precision}\n{engine}\n1\n")
    assert out_d == out_h
    assert tail_d.split("wall:")[0] == tail_h.split("wall:")[0]
    if dt0 > 1e-3:
        assert any("WARNING: time step rejected!" in ln for ln in out_d)
    got = {}
    for mode in ("host", "device"):
        got[mode] = np.fromfile(tmp_path / mode / "output" / "final_state.bin", dtype=np.float64)
        snaps = sorted(x for x in os.listdir(tmp_path / mode / "output") if x[0].isdigit())
        got[mode + "_snaps"] = snaps
    assert got["device_snaps"] == got["host_snaps"] and len(got["host_snaps"]) == 11
    assert np.array_equal(got["device"].view(np.uint64), got["host"].view(np.uint64))
    if precision == 0 and dt0 < 1e-3:
        cfg = oracle.Config(nx=n, ny=n, limiter=4, bc=(4, 4, 4, 4))
        ts = oracle.TimeSettings(t_final=t_final, cfl_target=0.9, dt_max=1e-2, n_snap=10)
        q = oracle.dam_break(cfg)
        oracle.run(cfg, ts, q)
        from tests.helpers import sides_mask

        m = sides_mask(n, n)
        assert np.array_equal(got["device"].reshape(n + 4, n + 4, 3)[m], q[m])
