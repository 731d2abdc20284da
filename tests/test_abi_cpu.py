"""CPU checks of the C ABI: the HIP library loads without a GPU, exports every
entry point include/pcw.h declares, and the ctypes mirrors of the structs
have the C layout (checked with a gcc-compiled probe of the header)."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pcw.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(pcw_\w+)\s*\(", text, re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    from paraclaws_amd import abi

    L = abi.load()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(abi.SIGNATURES), set(names) ^ set(abi.SIGNATURES)


def test_struct_layouts_match_header(tmp_path):
    from paraclaws_amd import abi

    probe = tmp_path / "probe.c"
    probe.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "pcw.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu\\n\","
        "sizeof(pcw_domain), sizeof(pcw_params), sizeof(pcw_time), sizeof(pcw_step_record),"
        "sizeof(pcw_status), offsetof(pcw_params, grav), offsetof(pcw_domain, xi_min),"
        "offsetof(pcw_time, n_snapshots));return 0;}\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(probe), "-o", str(exe)], check=True)
    got = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()))
    want = [C.sizeof(abi.PcwDomain), C.sizeof(abi.PcwParams), C.sizeof(abi.PcwTime),
            C.sizeof(abi.PcwStepRecord), C.sizeof(abi.PcwStatus), abi.PcwParams.grav.offset,
            abi.PcwDomain.xi_min.offset, abi.PcwTime.n_snapshots.offset]
    assert got == want


def test_create_rejects_bad_arguments_before_touching_the_gpu():
    from paraclaws_amd import abi

    with pytest.raises(abi.PcwError) as e:
        abi.Engine([abi.Domain(16, 16, bc=(1, 2, 2, 2))])
    assert e.value.code == abi.PCW_ERR_BC
    with pytest.raises(abi.PcwError) as e:
        abi.Engine([abi.Domain(16, 16)], precision=abi.FP32, engine=abi.STAGED)
    assert e.value.code == abi.PCW_ERR_ARG
    with pytest.raises(abi.PcwError) as e:
        abi.Engine([abi.Domain(1, 16)])
    assert e.value.code == abi.PCW_ERR_ARG
    with pytest.raises(abi.PcwError) as e:
        abi.Engine([abi.Domain(16, 16)], limiter=7)
    assert e.value.code == abi.PCW_ERR_ARG


def test_create_rejects_sizes_beyond_the_limits_before_touching_the_gpu():
    """nx, ny above PCW_MAX_SIDE and more than PCW_MAX_DOMAINS instances are
    refused with a message (the values match the header's)."""
    from paraclaws_amd import abi

    text = open(HEADER).read()
    assert int(re.search(r"#define PCW_MAX_SIDE (\d+)", text).group(1)) == abi.PCW_MAX_SIDE
    assert int(re.search(r"#define PCW_MAX_DOMAINS (\d+)", text).group(1)) == abi.PCW_MAX_DOMAINS
    for dom in (abi.Domain(abi.PCW_MAX_SIDE + 1, 16), abi.Domain(16, abi.PCW_MAX_SIDE + 1)):
        for prec in (abi.FP64, abi.FP32):
            with pytest.raises(abi.PcwError, match="PCW_MAX_SIDE") as e:
                abi.Engine([abi.Domain(16, 16), dom], precision=prec)
            assert e.value.code == abi.PCW_ERR_ARG
    with pytest.raises(abi.PcwError, match="PCW_MAX_DOMAINS") as e:
        abi.Engine([abi.Domain(4, 4)] * (abi.PCW_MAX_DOMAINS + 1))
    assert e.value.code == abi.PCW_ERR_ARG


def test_missing_library_fails_loudly(tmp_path):
    from paraclaws_amd import abi

    saved = abi._lib
    abi._lib = None
    try:
        with pytest.raises(RuntimeError, match="no CPU fallback"):
            abi.load(str(tmp_path / "nope.so"))
    finally:
        abi._lib = saved


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus 2 exits with status 2 and says why,
    before anything touches a GPU."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_step_noalloc_rejects_bad_arguments_before_touching_the_gpu():
    """pcw_step_noalloc checks its arguments before any HIP call (the device
    addresses below are never dereferenced) and reports through
    pcw_last_error(NULL)."""
    from paraclaws_amd import abi

    p = abi.make_params(abi.FP32)
    rc, _, msg = abi.step_noalloc(p, abi.Domain(64, 64, source=abi.SRC_BATHY), 0x10000, 0x900000, 1e-3)
    assert rc == abi.PCW_ERR_ARG and "bathymetry" in msg
    rc, _, msg = abi.step_noalloc(p, abi.Domain(64, 64), 0x10010, 0x900000, 1e-3)
    assert rc == abi.PCW_ERR_ARG and "aligned" in msg
    rc, _, msg = abi.step_noalloc(p, abi.Domain(64, 64), 0x10000, 0x10100, 1e-3)
    assert rc == abi.PCW_ERR_ARG and "overlap" in msg
    rc, _, msg = abi.step_noalloc(p, abi.Domain(64, 64, bc=(1, 2, 2, 2)), 0x10000, 0x900000, 1e-3)
    assert rc == abi.PCW_ERR_BC
    rc, _, msg = abi.step_noalloc(p, abi.Domain(64, abi.PCW_MAX_SIDE + 1), 0x10000, 0x900000, 1e-3)
    assert rc == abi.PCW_ERR_ARG and "PCW_MAX_SIDE" in msg
    rc, _, msg = abi.step_noalloc(abi.make_params(abi.FP64, tile_rows=2), abi.Domain(60 * 600, 1000),
                                  0x10000000, 0x80000000000, 1e-3)
    assert rc == abi.PCW_ERR_ARG and "tiles" in msg
    rc, _, msg = abi.step_noalloc(abi.make_params(abi.FP64, engine=abi.STAGED), abi.Domain(64, 64),
                                  0x10000, 0x900000, 1e-3)
    assert rc == abi.PCW_ERR_ARG


def test_frame_layout_round_trip():
    from paraclaws_amd import abi

    rng = np.random.default_rng(0)
    for prec in (abi.FP64, abi.FP32):
        for nx, ny in ((2, 3), (60, 5), (61, 4), (121, 7)):
            q = rng.normal(size=(ny + 4, nx + 4, 3)).astype(np.float32).astype(np.float64)
            planes = abi.frame_to_planes(q, prec)
            pitch, pos, _ = abi.frame_layout(prec, nx, ny)
            assert len(set(pos.tolist())) == nx + 4 and pos.max() < pitch
            assert np.array_equal(abi.planes_to_frame(planes, prec, nx, ny), q)
