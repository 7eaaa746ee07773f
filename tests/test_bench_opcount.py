"""The operation count bench.py reports beside the issue rate: the bitonic
network's stages and compare-exchanges per (wavenumber, layer) item, checked
against an explicit enumeration of the network's stages (every merge level k
of a 2^n-slot network has k stages; the levels within a lane's sorted run of
E slots are skipped)."""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))

import bench  # noqa: E402


def _stages(slots, per_lane):
    n = slots.bit_length() - 1
    skip = per_lane.bit_length() - 1
    return sum(k for k in range(1, n + 1) if k > skip)


def test_network_op_count_default_workload():
    ops = bench.network_op_count(20, 3)
    assert ops["slots"] == 512
    assert ops["network_stages_per_round"] == _stages(512, 8) == 39
    assert ops["compare_exchanges_per_item"] == 2 * 39 * 256
    assert ops["minmax_wave_instructions_per_item"] == 2 * 2 * 39 * 256 / 64


def test_network_op_count_other_shapes():
    # 16 ordinates: 4 slots per lane; 24: 16 per lane; one k-gas: no round
    assert bench.network_op_count(16, 2)["network_stages_per_round"] == _stages(256, 4)
    assert bench.network_op_count(24, 3)["network_stages_per_round"] == _stages(1024, 16)
    assert bench.network_op_count(20, 1)["compare_exchanges_per_item"] == 0
