"""Host logic of bench.py that runs without a GPU: the canonical byte models
the roofline figures divide by, the one-GPU contract, and the committed
profile summaries the benchmark line reads."""
import importlib.util
import json
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_fp64_model_counts_words_per_stage(bench):
    # 39 step-glue words, 27 per advected field (30 with the scalar update),
    # 37 per BiCGSTAB iteration, each 8 bytes
    assert bench.canonical_bytes_fp64(0, 0) == 8 * 39
    assert bench.canonical_bytes_fp64(10, 0) - bench.canonical_bytes_fp64(9, 0) == 8 * 27
    assert (bench.canonical_bytes_fp64(10, 0, scalar_update=True)
            - bench.canonical_bytes_fp64(9, 0, scalar_update=True)) == 8 * 30
    assert bench.canonical_bytes_fp64(10, 5) - bench.canonical_bytes_fp64(10, 4) == 8 * 37


def test_mixed_model_grows_with_fields_iterations_and_refinements(bench):
    base = bench.canonical_bytes_mixed(10, 4, 2, multigrid=True)
    # an MPDATA field: min (8) + donor (24) + antidiffusive velocities (28) +
    # limiter (32) + limited donor pass (32) bytes
    assert bench.canonical_bytes_mixed(11, 4, 2, multigrid=True) - base == pytest.approx(124)
    # an inner iteration of the multigrid path: MG_WORDS_PER_ITER FP32 words
    assert (bench.canonical_bytes_mixed(10, 5, 2, multigrid=True) - base
            == pytest.approx(4 * bench.MG_WORDS_PER_ITER))
    # a refinement: true residual, inner start, correction
    assert bench.canonical_bytes_mixed(10, 4, 3, multigrid=True) - base == pytest.approx(64)
    # the sweeps path: 37 words per iteration with four divisions, 3 more
    # per sweep in each of the two applications for every further division
    sweeps4 = bench.canonical_bytes_mixed(10, 1, 2) - bench.canonical_bytes_mixed(10, 0, 2)
    sweeps12 = (bench.canonical_bytes_mixed(10, 1, 2, divisions=12)
                - bench.canonical_bytes_mixed(10, 0, 2, divisions=12))
    assert sweeps4 == pytest.approx(4 * 37)
    assert sweeps12 == pytest.approx(4 * (37 + 6 * 8))
    assert bench.jac_per_iter(4) == 3 and bench.jac_per_iter(12) == 19


def test_rel_l2_removes_the_mean_when_asked(bench):
    a = np.random.default_rng(0).standard_normal((6, 7, 8))
    assert bench.rel_l2(a, a + 5.0, center=True) < 1e-15
    assert bench.rel_l2(a, a + 5.0) > 0.1


def test_more_than_one_gpu_is_refused(bench, capsys):
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err


def test_committed_profile_summaries_are_readable(bench):
    busy = bench.load_busy("moist_bubble_256x256x128_mixed")
    assert busy is not None and busy["kernel_ms_per_step"] > 0
    with open(os.path.join(ROOT, busy["source"])) as f:
        d = json.load(f)
    # the union of the kernels' intervals cannot exceed their summed durations
    assert d["kernel_ms_per_step"] <= d["kernel_sum_ms_per_step"]
    traffic = bench.load_traffic("moist_bubble_256x256x128_mixed", kernel="k32_betalim")
    assert traffic is not None and traffic > 0
