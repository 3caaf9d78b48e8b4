"""Measurement tooling on the CPU: the rocprofv3 summaries that the profile
files are written with, on small synthetic result databases, and the bench's
refusal of more than one GPU."""
import os
import sqlite3
import subprocess
import sys

import pytest

from barbu_amd import rocprof_summary as rs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NS = "(anonymous namespace)::"


def kernels_db(path, rows):
    con = sqlite3.connect(path)
    con.execute("create table kernels (name text, kernel_id int, vgpr_count int, lds_size int, duration real)")
    con.executemany("insert into kernels values (?, ?, ?, ?, ?)", rows)
    con.commit()
    con.close()


def counters_db(path, rows):
    con = sqlite3.connect(path)
    con.execute("create table counters_collection (kernel_name text, kernel_id int, grid_size int, "
                "workgroup_size int, vgpr_count int, counter_name text, value real, dispatch_id int, duration real)")
    con.executemany("insert into counters_collection values (?, ?, ?, ?, ?, ?, ?, ?, ?)", rows)
    con.commit()
    con.close()


def test_kernel_names():
    assert rs.short("void " + NS + "ftl_tiles_kernel<true>(FtlStep const*, FtlSpan const*, int, int)") == \
        "ftl_tiles_kernel<true>"
    assert rs.short(NS + "ftl_plain_wind_kernel<false>(FtlStep)") == "ftl_plain_wind_kernel<false>"
    assert rs.short("ftl_copy_kernel(float4 const*, float4*, unsigned long)") == "ftl_copy_kernel"
    assert rs.base("ftl_tiles_kernel<true>") == "ftl_tiles_kernel"
    assert rs.base("ftl_copy_kernel") == "ftl_copy_kernel"


def test_trace_groups_by_name_and_object(tmp_path):
    db = str(tmp_path / "t.db")
    k = "void " + NS + "ftl_tiles_kernel<true>(FtlStep const*)"
    kernels_db(db, [(k, 3, 64, 8192, 100_000.0), (k, 3, 64, 8192, 300_000.0), (k, 5, 36, 4096, 50_000.0),
                    ("ftl_copy_kernel(void)", 9, 8, 0, 50_000.0)])
    lines = rs.trace(db).splitlines()
    row = [l.split() for l in lines if l.startswith("ftl_tiles_kernel<true>")][0]
    assert row[1] == "3" and float(row[3]) == pytest.approx(150.0)  # calls, avg_us (both objects)
    ids = [l.split() for l in rs.trace_ids(db).splitlines() if l.startswith("ftl_tiles_kernel<true>")]
    assert sorted((r[1], r[4], r[6]) for r in ids) == [("3", "2", "200.00"), ("5", "1", "50.00")]


def test_traffic_scaled_by_the_copy_kernel(tmp_path):
    """Bytes per dispatch = counter KiB x 1024 x (copy bytes / what the copy
    kernel's counter reads); kernels are picked by base name."""
    copy_bytes = 1 << 20
    fk, wk = str(tmp_path / "f.db"), str(tmp_path / "w.db")
    tiles = NS + "ftl_tiles_kernel<true>(FtlStep const*)"
    # the copy kernel's FETCH_SIZE reads half its bytes (x2 scale), WRITE_SIZE all of them (x1)
    counters_db(fk, [("ftl_copy_kernel()", 1, 4096, 256, 8, "FETCH_SIZE", 512.0, 1, 10.0),
                     (tiles, 2, 6400, 64, 64, "FETCH_SIZE", 1000.0, 2, 20.0),
                     (tiles, 2, 6400, 64, 64, "FETCH_SIZE", 1100.0, 3, 20.0),
                     (tiles, 4, 1280, 64, 36, "FETCH_SIZE", 100.0, 4, 5.0)])
    counters_db(wk, [("ftl_copy_kernel()", 7, 4096, 256, 8, "WRITE_SIZE", 1024.0, 1, 10.0),
                     (tiles, 8, 6400, 64, 64, "WRITE_SIZE", 2000.0, 2, 20.0),
                     (tiles, 9, 1280, 64, 36, "WRITE_SIZE", 300.0, 4, 5.0)])
    text, (rb, wb) = rs.traffic_ids(fk, wk, copy_bytes, {"ftl_tiles_kernel"})
    # launches matched across the runs by (kernel, workgroups, VGPRs), whatever their object ids
    assert rb == pytest.approx((1050.0 + 100.0) * 1024 * 2)
    assert wb == pytest.approx((2000.0 + 300.0) * 1024 * 1)
    assert "x2.0000" in text and "x1.0000" in text
    _, res = rs.traffic(fk, wk, copy_bytes)
    assert res["ftl_tiles_kernel<true>"][1] == pytest.approx((2000.0 + 300.0) / 2 * 1024)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
