"""Uneven-batch host logic without a device: job packing and the checks
cfft_uneven_create makes before touching the GPU."""
import numpy as np
import pytest


def test_packed_jobs_layout():
    from cfft_amd.uneven import packed_jobs
    jobs, nin, nf, nout = packed_jobs([9, 5, 14])
    assert list(jobs["in_offset"]) == [0, 512, 544]
    assert list(jobs["fourier_offset"]) == [0, 256, 272]
    assert nin == nout == 512 + 32 + 16384 and nf == nin // 2


@pytest.mark.parametrize("log2_N,status", [(4, 2), (15, 2)])
def test_create_rejects_sizes(log2_N, status):
    from cfft_amd import CfftPanic
    from cfft_amd.uneven import Schedule, packed_jobs
    jobs, nin, nf, nout = packed_jobs([10, log2_N])
    with pytest.raises(CfftPanic) as ei:
        Schedule(jobs, nin, nf, nout)
    assert ei.value.status == status and "log2_N" in str(ei.value)


def test_create_rejects_out_of_bounds_jobs():
    from cfft_amd import CfftPanic
    from cfft_amd.uneven import Schedule, packed_jobs
    jobs, nin, nf, nout = packed_jobs([10, 11, 12])
    for field, length in (("in_offset", nin), ("fourier_offset", nf), ("out_offset", nout)):
        bad = jobs.copy()
        bad[field][2] += 1
        lens = [nin, nf, nout]
        with pytest.raises(CfftPanic, match="job 2: buffer range") as ei:
            Schedule(bad, *lens)
        assert ei.value.status == 10
    with pytest.raises(CfftPanic, match="job 0"):
        Schedule(jobs, nin, nf, 100)
