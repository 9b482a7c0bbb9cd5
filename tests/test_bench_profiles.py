"""bench.py reads the measured HBM traffic and VALU instruction count of the step kernel from
the committed profile summaries, only for the tile shape they were measured on."""
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import bench  # noqa: E402


def test_profile_values_for_the_measured_tile():
    assert bench.read_traffic("step_c3_fp32_4096", (32, 32, 512)) == 136843871
    assert bench.read_valu("step_c3_fp32_4096", (32, 32, 512)) == 222.0
    # the traffic is within a few per cent of the algorithmic 8 B per site
    assert 1.0 <= bench.read_traffic("step_c3_fp32_4096", (32, 32, 512)) / (8 * 4096 ** 2) < 1.05


def test_profile_values_refused_for_other_tiles_or_missing_files():
    assert bench.read_traffic("step_c3_fp32_4096", (32, 48, 512)) is None
    assert bench.read_valu("step_c3_fp32_4096", (32, 48, 512)) is None
    assert bench.read_traffic("no_such_profile", (32, 32, 512)) is None
    assert bench.read_valu("", (32, 32, 512)) is None
