"""C++ wrapper (include/mtcheck.hpp) and the two file commands built on it."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import first_mismatch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "mtcheck_amd", "bin")


def test_cpp_wrapper_compiles_with_gcc(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "mtcheck.hpp"\nint main(){ mtcheck::TreeDeduplicator* p = nullptr; (void)p; '
                   'mtcheck::BasicRestarter* q = nullptr; (void)q; return mtcheck::get_mode(0, nullptr); }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_commands_print_usage_without_arguments():
    for prog in ("dedup_chkpt_files", "restart_chkpt_files"):
        r = subprocess.run([os.path.join(BIN, prog)], capture_output=True, text=True)
        assert r.returncode == 1 and "usage" in r.stderr


@pytest.mark.gpu
def test_file_commands_round_trip(tmp_path):
    """dedup_chkpt_files writes <file>.hashtree.incr_chkpt identical to the oracle's
    diffs and the reference's three CSV logs; restart_chkpt_files rebuilds each
    checkpoint byte for byte."""
    from mtcheck_amd.workload import CheckpointSeries
    C, nbytes, k = 128, (1 << 20) + 77, 4
    series = CheckpointSeries(nbytes, C, seed=5)
    buf = np.zeros(nbytes, np.uint8)
    orc = O.OracleDedup(C)
    files, expect = [], []
    for i in range(k):
        series.apply_host(buf, i)
        path = tmp_path / f"chk{i}.dat"
        path.write_bytes(buf.tobytes())
        files.append(str(path))
        expect.append(orc.checkpoint(buf, make_baseline=(i == 0)))
    print("dedup_chkpt_files", flush=True)
    r = subprocess.run([os.path.join(BIN, "dedup_chkpt_files"), str(C), str(k), "--run-tree-low-offset-chkpt",
                        *files], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    for i in range(k):
        got = open(files[i] + ".hashtree.incr_chkpt", "rb").read()
        same = got == expect[i]  # compared outside the assert: a rendered diff of 1 MiB takes minutes
        assert same, (i, first_mismatch(got, expect[i]))
        log = (tmp_path / f"chk{i}.dat.chunk_size.{C}.csv").read_text().splitlines()
        assert log[0].startswith("Approach,Chkpt ID") and log[1].startswith(f"Tree,{i},{C},{nbytes},")
        row = log[1].split(",")
        # Comparison Time (labels and forests), Gather Time (the diff) and Write Time
        # (the copy to the host) all measured
        assert float(row[8]) > 0 and float(row[9]) > 0 and float(row[10]) > 0, row
        size = (tmp_path / f"chk{i}.dat.chunk_size.{C}.size.csv").read_text().strip().split(",")
        assert size[:3] == ["Tree", str(i), str(C)] and size[5] == "32"
    for rid in (0, k - 1):
        print(f"restart_chkpt_files {rid}", flush=True)
        r = subprocess.run([os.path.join(BIN, "restart_chkpt_files"), str(rid), str(k), "1", str(C),
                            "--run-tree-low-offset-chkpt", *files], capture_output=True, text=True, cwd=tmp_path,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        lines = [l for l in r.stdout.splitlines() if "digest" in l]
        assert lines[0].split(":")[1].strip() == lines[1].split(":")[1].strip()
        assert (tmp_path / f"chk{rid}.dat.chunk_size.{C}.restart_timing.csv").exists()


@pytest.mark.gpu
def test_restart_command_basic_mode(tmp_path):
    """restart_chkpt_files --run-basic-chkpt rebuilds checkpoints from <file>.basic.incr_chkpt
    chains (written here by the oracle's basic approach) and logs 'Basic' restart rows."""
    from mtcheck_amd.workload import CheckpointSeries
    C, nbytes, k = 64, (1 << 19) + 5, 4
    series = CheckpointSeries(nbytes, C, seed=8)
    buf = np.zeros(nbytes, np.uint8)
    orc = O.OracleBasic(C)
    files = []
    for i in range(k):
        series.apply_host(buf, i)
        path = tmp_path / f"chk{i}.dat"
        path.write_bytes(buf.tobytes())
        files.append(str(path))
        (tmp_path / f"chk{i}.dat.basic.incr_chkpt").write_bytes(orc.checkpoint(buf, make_baseline=(i == 0)))
    for rid in (0, k - 1):
        print(f"restart_chkpt_files basic {rid}", flush=True)
        r = subprocess.run([os.path.join(BIN, "restart_chkpt_files"), str(rid), str(k), "1", str(C),
                            "--run-basic-chkpt", *files], capture_output=True, text=True, cwd=tmp_path, timeout=120)
        assert r.returncode == 0, r.stderr
        lines = [l for l in r.stdout.splitlines() if "digest" in l]
        assert lines[0].startswith("Basic chkpt digest")
        assert lines[0].split(":")[1].strip() == lines[1].split(":")[1].strip()
        log = (tmp_path / f"chk{rid}.dat.chunk_size.{C}.restart_timing.csv").read_text().splitlines()
        assert log[-1].startswith(f"Basic,{rid},{C},")


@pytest.mark.gpu
def test_file_commands_basic_mode(tmp_path):
    """dedup_chkpt_files --run-basic-chkpt writes <file>.basic.incr_chkpt identical to the
    oracle's basic diffs and logs 'Basic' rows (the size log counting the checkpoints
    since the baseline, src/basic_approach.cpp:672); restart_chkpt_files rebuilds each."""
    from mtcheck_amd.workload import CheckpointSeries
    C, nbytes, k = 64, (1 << 19) + 5, 4
    series = CheckpointSeries(nbytes, C, seed=9)
    buf = np.zeros(nbytes, np.uint8)
    orc = O.OracleBasic(C)
    files, expect = [], []
    for i in range(k):
        series.apply_host(buf, i)
        path = tmp_path / f"chk{i}.dat"
        path.write_bytes(buf.tobytes())
        files.append(str(path))
        expect.append(orc.checkpoint(buf, make_baseline=(i == 0)))
    print("dedup_chkpt_files basic", flush=True)
    r = subprocess.run([os.path.join(BIN, "dedup_chkpt_files"), str(C), str(k), "--run-basic-chkpt", *files],
                       capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    for i in range(k):
        got = open(files[i] + ".basic.incr_chkpt", "rb").read()
        same = got == expect[i]
        assert same, (i, first_mismatch(got, expect[i]))
        log = (tmp_path / f"chk{i}.dat.chunk_size.{C}.csv").read_text().splitlines()
        assert log[1].startswith(f"Basic,{i},{C},{nbytes},")
        size = (tmp_path / f"chk{i}.dat.chunk_size.{C}.size.csv").read_text().strip().split(",")
        assert size[:3] == ["Basic", str(i), str(C)] and len(size) == 9 + i, size
    for rid in (0, k - 1):
        print(f"restart_chkpt_files basic {rid}", flush=True)
        r = subprocess.run([os.path.join(BIN, "restart_chkpt_files"), str(rid), str(k), "1", str(C),
                            "--run-basic-chkpt", *files], capture_output=True, text=True, cwd=tmp_path, timeout=120)
        assert r.returncode == 0, r.stderr
        lines = [l for l in r.stdout.splitlines() if "digest" in l]
        assert lines[0].split(":")[1].strip() == lines[1].split(":")[1].strip()
