"""Point-file format (common.h:245-266: size_t count, then raw elements) and
the command-line driver with the reference driver's interface
(testing/floatN-knn-and-fcp.cu:514-558)."""
import os
import re
import subprocess

import numpy as np
import pytest

import cudakdtree_amd as ck
from cudakdtree_amd import datagen, io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cudakdtree_amd", "cli", "cukd_knn_fcp")


def test_point_file_round_trip_and_layout(tmp_path):
    pts = datagen.uniform(1234, 3, 5)
    f = tmp_path / "pts"
    io.save_points(str(f), pts)
    raw = f.read_bytes()
    assert len(raw) == 8 + pts.nbytes
    assert int(np.frombuffer(raw[:8], np.uint64)[0]) == 1234  # little-endian size_t count
    assert raw[8:] == pts.tobytes()
    assert np.array_equal(io.load_points(str(f), 3), pts)
    rec = np.dtype([("p", np.float32, 3), ("dim", np.int32)])
    a = np.zeros(10, rec)
    a["p"] = datagen.uniform(10, 3, 1)
    io.save_points(str(f), a)
    assert np.array_equal(io.load_points(str(f), dtype=rec), a)


def test_point_file_truncated(tmp_path):
    f = tmp_path / "bad"
    f.write_bytes(np.uint64(100).tobytes() + b"\0" * 12)
    with pytest.raises(ValueError):
        io.load_points(str(f), 3)


def test_cli_is_built_and_rejects_bad_args():
    assert os.path.exists(CLI), "run __graft_entry__.build()"
    r = subprocess.run([CLI, "--no-such-flag"], capture_output=True, text=True)
    assert r.returncode == 1 and "unknown cmdline arg" in r.stderr
    r = subprocess.run([CLI, "--dims", "5"], capture_output=True, text=True)
    assert r.returncode == 1


def _run(args, cwd):
    r = subprocess.run([CLI] + args, capture_output=True, text=True, cwd=cwd, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _checksum(out):
    return float(re.search(r"CHECKSUM (\S+)", out).group(1))


@pytest.mark.gpu
@pytest.mark.parametrize("method,extra", [
    ("stackBased", []), ("stackFree", []), ("cct", []), ("stackBased", ["--xd"]),
    ("cct", ["--xd", "--clustered"]), ("spatial-stackBased", ["-lt", "4"]), ("spatial-cct", []),
])
@pytest.mark.parametrize("mode", [[], ["-k", "8"], ["-k", "4", "-r", "30000"]])
def test_cli_verifies(method, extra, mode, tmp_path):
    out = _run(["20000", "-nq", "3000", "-nr", "2", "-v", "--stats", "--method", method]
               + extra + mode, tmp_path)
    assert "verification succeeded" in out and "that is" in out
    if not method.startswith("spatial"):
        assert "valid k-d tree" in out
    assert re.search(r"NICE_STATS \S+", out)


@pytest.mark.gpu
def test_cli_dumped_files_round_trip(tmp_path):
    a = _run(["5000", "-nq", "700", "-k", "8", "--dims", "4", "--dump-files"], tmp_path)
    assert (tmp_path / "data_points").exists() and (tmp_path / "query_points").exists()
    pts = io.load_points(str(tmp_path / "data_points"), 4)
    assert pts.shape == (5000, 4) and np.all(pts == np.round(pts)) and pts.max() <= 1e6
    b = _run(["--load-dumped-files", "-k", "8", "--dims", "4"], tmp_path)
    assert _checksum(a) == _checksum(b)
