"""CPU checks of bench.py's host logic: the lookup roofline's count of the
scattering lookups a batch makes, the rocprofv3 summaries of profiles/, and
the one-GPU guard."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

RB, RT = 6360.0, 6420.0


def rays(cams, views):
    f = lambda a: np.asarray(a, dtype=np.float32)
    return dict(camera=f(cams), view_ray=f(views))


def test_scattering_lookups_per_ray():
    # Inside the atmosphere: GetSkyRadiance 1 lookup, ToPoint 2.
    assert bench.scattering_lookups(rays([[0, 0, RB + 10]], [[1, 0, 0]]), RT) == 3.0
    # Above it, looking down: both GetSkyRadiance's lookup and ToPoint's
    # camera-side lookup are made at the top radius (8 distinct texels of
    # 16); ToPoint's second lookup is inside.
    assert bench.scattering_lookups(rays([[0, 0, RT + 30]], [[0, 0, -1]]), RT) == 2.0
    # Above it, looking away: GetSkyRadiance makes none.
    assert bench.scattering_lookups(rays([[0, 0, RT + 30]], [[0, 0, 1]]), RT) == 1.5
    # Above it, a ray that passes outside the top sphere.
    assert bench.scattering_lookups(rays([[0, 0, RT + 30]], [[1, 0, 0]]), RT) == 1.5


def test_scattering_lookups_of_the_workload():
    from precomputed_atmospheric_scattering_amd import configs
    q = configs.sky_queries(2024, 20000, RB, RT)
    per_ray = bench.scattering_lookups(q, RT) / 20000
    # 40% of the cameras are above the atmosphere (altitude U[0, 100] km).
    assert 2.4 < per_ray < 2.6


def write_csv(path, header, rows):
    with open(path, "w") as f:
        f.write(",".join('"%s"' % h for h in header) + "\n")
        for r in rows:
            f.write(",".join(str(x) for x in r) + "\n")


def test_summaries(tmp_path):
    stats = tmp_path / "trace"
    stats.mkdir()
    write_csv(stats / "run_kernel_stats.csv",
              ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs",
               "StdDev"],
              [['"k_a"', 2, 4000000, 2000000, 80.0, 1900000, 2100000, 0],
               ['"__amd_rocclr_copyBuffer"', 1, 1000000, 1000000, 20.0, 1000000, 1000000, 0]])
    table = bench.summarize_stats(str(stats))
    assert "| k_a | 2 | 2.000 | 1.900 | 2.100 | 4.00 | 80.0 |" in table
    cols = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    sq, fe, wr = (tmp_path / n for n in ("sq", "fetch", "write"))
    for d in (sq, fe, wr):
        d.mkdir()
    write_csv(sq / "run_counter_collection.csv", cols,
              [[1, '"void atmo::k_a<4>(int)"', "SQ_INSTS_VALU", 1e9],
               [1, '"void atmo::k_a<4>(int)"', "SQ_WAVES", 1000]])
    # The copy probe reads and writes 256 MiB: FETCH_SIZE counts half of it.
    write_csv(fe / "run_counter_collection.csv", cols,
              [[1, '"__amd_rocclr_copyBuffer"', "FETCH_SIZE", 128 * 1024],
               [2, '"void atmo::k_a<4>(int)"', "FETCH_SIZE", 1000]])
    write_csv(wr / "run_counter_collection.csv", cols,
              [[1, '"__amd_rocclr_copyBuffer"', "WRITE_SIZE", 256 * 1024],
               [2, '"void atmo::k_a<4>(int)"', "WRITE_SIZE", 500]])
    table = bench.summarize_counters(str(stats), str(sq), str(fe), str(wr))
    assert "FETCH_SIZE x 2.000, WRITE_SIZE x 1.000" in table
    row = [l for l in table.splitlines() if l.startswith("| k_a |")][0].split("|")
    # 1e9 wave instructions x 64 lanes over 2 ms; 2000 KiB fetched (x2),
    # 500 KiB written.
    assert row[6].strip() == "1000000"
    assert row[10].strip() == "2.0" and row[11].strip() == "0.5"
    assert row[13].strip() == "32.00"


def test_valu_mix_summary(tmp_path):
    trace, m32, m64 = (tmp_path / n for n in ("trace", "m32", "m64"))
    for d in (trace, m32, m64):
        d.mkdir()
    write_csv(trace / "run_kernel_stats.csv",
              ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs",
               "StdDev"],
              [['"k_a"', 1, 1000000, 1000000, 100.0, 1000000, 1000000, 0]])
    cols = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    k = '"void atmo::k_a<4>(int)"'
    write_csv(m32 / "run_counter_collection.csv", cols,
              [[1, k, "SQ_INSTS_VALU", 1000], [1, k, "SQ_INSTS_VALU_FMA_F32", 500],
               [1, k, "SQ_INSTS_VALU_ADD_F32", 100], [1, k, "SQ_INSTS_VALU_MUL_F32", 100],
               [1, k, "SQ_INSTS_VALU_CVT", 50], [1, k, "SQ_INSTS_VALU_INT32", 150]])
    write_csv(m64 / "run_counter_collection.csv", cols,
              [[1, k, "SQ_INSTS_VALU", 1000], [1, k, "SQ_INSTS_VALU_FMA_F64", 100]])
    table = bench.summarize_valu_mix(str(trace), str(m32), str(m64))
    row = [l for l in table.splitlines() if l.startswith("| k_a |")][0].split("|")
    # 700 FP32 + 100 FP64 arithmetic wave instructions x 64 lanes in 1 ms.
    assert row[5].strip() == "700" and row[6].strip() == "100"
    assert row[7].strip() == "70%" and row[10].strip() == "5%" and row[11].strip() == "15%"
    rate = 800 * 64 / 1e-3 / 1e12
    assert row[12].strip() == "%.2f" % rate
    assert row[13].strip() == "%.2f" % (rate / bench.PEAK_FP_ARITH_TOPS)


def test_fp_arith_peak():
    # One FP arithmetic instruction per SIMD per 4 clocks: 256 CUs x 4 SIMDs
    # x 64 lanes / 4 clocks x 2.4 GHz.
    assert abs(bench.PEAK_FP_ARITH_TOPS - 39.3216) < 1e-9


def test_more_than_one_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
