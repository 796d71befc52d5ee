"""CPU checks of bench.py's host side: the one-GPU refusal and the argument checks (nothing
here touches a GPU), the algorithmic bytes per cell-step, the measured HBM traffic it reads
from profiles/ for the `roofline.traffic` field, and the fields of the cpu_baseline leg on a
small Kelvin-Helmholtz sample."""
import numpy as np
import pytest

import bench
from fv2d_amd import problems as P


def test_refuses_more_than_one_gpu(capsys):
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err


@pytest.mark.parametrize("argv", [["--steps", "0"], ["--warmup", "-1"], ["--serve-steps", "0"],
                                  ["--serve-jobs", "0"]])
def test_rejects_bad_counts(argv, capsys):
    assert bench.main(argv) == 2
    assert "must be" in capsys.readouterr().err


def test_algorithmic_bytes_per_cell_step():
    # an RK2 step reads U^n (stage 1), U* and U^n (stage 2) and writes U*, U^{n+1}: 5 states
    # of 4 fields; FP64 8 B, FP32 state 4 B, the FP32 form with the FP64 state 8 B
    assert bench.BYTES_PER_CELL_STEP == {"fp64": 5 * 4 * 8.0, "fp32": 5 * 4 * 4.0, "fp32s64": 5 * 4 * 8.0}
    assert bench.STAGE_LAUNCHES_PER_STEP == 2


@pytest.mark.parametrize("prec,bytes_per_cell_launch", [("fp64", 80), ("fp32", 40)])
def test_measured_traffic_from_profiles(prec, bytes_per_cell_launch):
    t = bench.read_traffic(prec, 4096)
    assert t is not None
    ratio = t / (bytes_per_cell_launch * 4096 * 4096)
    assert 1.0 <= ratio <= 1.5  # measured HBM bytes at or a little above the algorithmic ones
    assert bench.read_traffic(prec, 123) is None  # no profile for that size


def test_cpu_baseline_fields():
    size = 64
    q0 = P.kh_ic(P.kh_params(N=size))
    r = bench.run_cpu(size, 0.05, q0, samples=3)
    assert r["value"] > 0 and len(r["samples"]) == 3 and len(r["steps"]) == 3
    assert abs(r["value"] - float(np.median(r["samples"]))) <= 1e-5  # samples are rounded to 5 places
    assert r["threads"] >= 1 and r["stream_triad_gbs"] > 0
    for key in ("nproc", "cpu_model", "OMP_NUM_THREADS"):
        assert key in r
