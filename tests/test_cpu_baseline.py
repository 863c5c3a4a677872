"""The CPU baselines bench.py reports beside the GPU figure (CPU only).

* one core: the compiled reference's flux_div on a sample of the TGV workload;
* every core: the same driver as one single-rank process per core, run at
  once; the aggregate is the sum of the per-process rates, so with two
  processes it lies between one and two single-process rates (allowing for
  the noise of a shared host).
Both need the compiled reference (oracle/_ref), which is built only where the
reference sources are present.
"""
import os

import pytest

import bench

REF = os.path.join(bench.ROOT, "oracle", "_ref", "ref_driver")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="compiled reference (oracle/_ref) not built")


def test_one_core_reference_baseline():
    cb = bench.cpu_baseline(sample_blocks=1, reps=3)
    assert cb["kind"] == "reference" and cb["cores"] == 1
    assert cb["value"] > 0.0


def test_all_core_reference_baseline_sums_processes():
    one = bench.cpu_baseline(sample_blocks=1, reps=3)["value"]
    allc = bench.cpu_baseline_all_cores(sample_blocks=1, reps=3, max_procs=2)
    assert allc is not None and allc["kind"] == "reference"
    assert allc["cores"] == min(2, len(os.sched_getaffinity(0)))
    assert 0.5 * one < allc["value"] < 2.0 * allc["cores"] * one
