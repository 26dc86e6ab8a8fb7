"""The bench line's roofline fields come from the committed rocprofv3
summaries under profiles/: every (workload, precision) the bench times has a
summary at its launch shape, and the numbers the bench reads from it parse
and agree with the summary's own derived figures.  CPU only: no kernel runs."""
import os
import re
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

BEAMS = {"ma1000x400": 1000 * 400, "maa256x64x200": 256 * 64 * 200, "ensemble96": None}


def _launch(shape, grid, block):
    mode, cpt, bpt, blk = shape
    return {"mode": mode, "cpt": cpt, "beams_per_tile": bpt, "block": blk or block, "grid": grid}


def _dispatch_shape(text):
    """Grid and block of the profiled stepper dispatch, as its summary states them."""
    m = re.search(r"([0-9,]+) workgroups x\s+([0-9]+) threads", text)
    return (int(m.group(1).replace(",", "")), int(m.group(2))) if m else (None, None)


ENTRIES = [(key, i) for key in sorted(bench.PROFILES) for i in range(len(bench.PROFILES[key]))]


@pytest.mark.parametrize("key,i", ENTRIES)
def test_profile_summaries_feed_the_bench_line(key, i):
    workload, prec = key
    shape, hbm, valu = bench.PROFILES[key][i]
    for f in (hbm, valu):
        assert os.path.exists(os.path.join(bench.PROF, f)), f
    text = open(os.path.join(bench.PROF, valu)).read() + open(os.path.join(bench.PROF, hbm)).read()
    grid, block = _dispatch_shape(text)
    if grid is None:                     # the ensemble's summary gives the block only
        grid, block = 256, shape[3]
    launch = _launch(shape, grid, block)
    assert bench._profile(workload, prec, launch) == (os.path.join(bench.PROF, hbm), os.path.join(bench.PROF, valu))

    steps, kernel_ms = 100, 2.0
    cells = BEAMS[workload] or 14_200_000
    roof = bench._roofline(workload, prec, cells, steps, kernel_ms, launch)
    assert roof["traffic"] is not None and roof["traffic"] > 0, "traffic_bytes_per_step did not parse"
    per_step = roof["traffic"] / steps
    assert per_step == bench._profile_value(os.path.join(bench.PROF, hbm), r"^traffic_bytes_per_step:\s*([0-9.eE+]+)")
    # the measured bytes are the algorithmic ones scaled by the stated ratio
    assert abs(roof["traffic_vs_algorithmic"] - per_step / (bench._state_bytes(prec) * cells)) < 1e-12

    rv = bench._roofline_valu(workload, prec, steps, kernel_ms, launch)
    assert rv is not None, "SQ_INSTS_VALU row did not parse"
    assert 100 < rv["valu_per_wave_step"] < 1e6
    assert rv["waves"] == grid * (block // 64)
    assert 0 < rv["fp32_tflops"] + rv["fp64_tflops"]


def test_other_launch_shapes_get_no_profile_figures():
    """A launch shape the summaries were not taken at reports no traffic and
    no VALU roofline rather than another shape's numbers."""
    shape, _, _ = bench.PROFILES[("ma1000x400", "fp32")][0]
    other = _launch((shape[0], 2 if shape[1] != 2 else 1, shape[2], shape[3]), 250, 448)
    assert bench._profile("ma1000x400", "fp32", other) is None
    roof = bench._roofline("ma1000x400", "fp32", 400000, 10, 1.0, other)
    assert roof["traffic"] is None
    assert bench._roofline_valu("ma1000x400", "fp32", 10, 1.0, other) is None
