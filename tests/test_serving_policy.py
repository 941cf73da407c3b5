"""Admission order of the continuous-batching server (host logic, no GPU):
arrival order under "fifo"; under "sjf" the smallest work (voxel-steps)
first, except that jobs waiting longer than max_wait go first in arrival
order."""
import numpy as np
import pytest

from bpm_amd.serving import BpmServer, Job, job_work


def _job(jid, arrival, N, nz, n_angles):
    return Job(jid, arrival, N, nz, np.zeros((n_angles, 2)), 0)


def test_job_work_counts_voxel_steps():
    assert job_work(_job(0, 0., 128, 65, 8)) == 128 * 128 * 64 * 8


def test_fifo_and_sjf_orders():
    q = [_job(0, 0.0, 1024, 512, 64), _job(1, 0.1, 128, 64, 8), _job(2, 0.2, 512, 100, 10),
         _job(3, 0.3, 256, 64, 8)]
    fifo = BpmServer(1 << 30, policy="fifo")
    assert [j.jid for j in fifo._order(q, now=0.5)] == [0, 1, 2, 3]
    sjf = BpmServer(1 << 30, policy="sjf", max_wait=1.0)
    # nobody has waited a second: smallest work first
    assert [j.jid for j in sjf._order(q, now=0.5)] == [1, 3, 2, 0]
    # job 0 (arrived at 0) and job 1 have now waited longer than max_wait
    assert [j.jid for j in sjf._order(q, now=1.15)] == [0, 1, 3, 2]
    with pytest.raises(ValueError):
        BpmServer(1 << 30, policy="lifo")
