"""Scheduling logic of SimServer on the CPU, against a simulated batch.

The fake batch has the FemBatch surface (footprint / free_bytes / add / remove /
step / sync / progress / positions / velocities) but only counts substeps, so
admission order, footprint waits, horizon retirement and failure isolation are
checked without a GPU.
"""

from types import SimpleNamespace

import numpy as np
import pytest

from flatworld_amd import server as server_mod
from flatworld_amd.server import Job, SimServer, default_prep_workers, poisson_arrivals


class FakeBatch:
    def __init__(self, capacity):
        self.capacity = capacity
        self.used = 0
        self.scenes = {}
        self.next_id = 0
        self.max_resident = 0
        self.admit_order = []
        self.frames = []

    def footprint(self, spec):
        return spec.bytes

    def free_bytes(self):
        return self.capacity - self.used

    def add(self, spec):
        if spec.bytes > self.free_bytes():
            raise RuntimeError("arena full")
        sid = self.next_id
        self.next_id += 1
        self.used += spec.bytes
        self.scenes[sid] = {"spec": spec, "ran": 0}
        self.max_resident = max(self.max_resident, len(self.scenes))
        self.admit_order.append(spec.name)
        return sid

    def remove(self, sid):
        self.used -= self.scenes.pop(sid)["spec"].bytes

    def step(self, n, graph=True, frame=None):
        self.frames.append(n)
        for s in self.scenes.values():
            s["ran"] += n

    def sync(self):
        pass

    def progress(self, sid):
        s = self.scenes[sid]
        done = min(max(s["ran"] - 1, 0), s["spec"].horizon)  # first substep is the warm-up
        return done, done >= s["spec"].horizon

    def positions(self, sid):
        s = self.scenes[sid]
        v = np.nan if s["spec"].poison else float(s["spec"].name)
        return np.full((4, 2), v, np.float32)

    def velocities(self, sid):
        return np.zeros((4, 2), np.float32)


def spec(name, horizon, nbytes, poison=False):
    return SimpleNamespace(name=name, horizon=horizon, bytes=nbytes, n_elems=10, poison=poison)


def test_fifo_admission_and_retirement():
    b = FakeBatch(capacity=100)
    srv = SimServer(batch=b, frame=60)
    jobs = [Job(i, spec(i, h, 40), 0.0) for i, h in enumerate([30, 200, 59, 60, 500])]
    res = srv.run(jobs)
    assert sorted(r.job_id for r in res) == list(range(5))
    assert b.admit_order == [0, 1, 2, 3, 4]  # arrival order
    assert b.max_resident == 2  # 100 bytes hold two 40-byte jobs
    assert all(r.ok for r in res)
    # horizon 59 needs warm-up + 59 substeps = 60: one frame once admitted
    by = {r.job_id: r for r in res}
    assert by[0].finished <= by[1].finished


def test_failed_job_is_isolated():
    b = FakeBatch(capacity=1000)
    srv = SimServer(batch=b, frame=60)
    jobs = [Job(0, spec(0, 60, 10)), Job(1, spec(1, 60, 10, poison=True)), Job(2, spec(2, 60, 10))]
    res = {r.job_id: r for r in srv.run(jobs)}
    assert not res[1].ok and res[0].ok and res[2].ok
    assert srv.stats()["failed"] == 1


def test_oversized_job_rejected():
    b = FakeBatch(capacity=10)
    srv = SimServer(batch=b, frame=60)
    with pytest.raises(RuntimeError, match="needs"):
        srv.run([Job(0, spec(0, 60, 11))])


def test_horizon_required():
    srv = SimServer(batch=FakeBatch(10), frame=60)
    with pytest.raises(ValueError, match="horizon"):
        srv.run([Job(0, spec(0, 0, 1))])


def test_poisson_arrivals_rate():
    t = poisson_arrivals(20000, 500.0, np.random.default_rng(0))
    assert np.all(np.diff(t) > 0)
    assert abs(20000 / t[-1] - 500.0) < 15.0


class FakePrepBatch(FakeBatch):
    """FakeBatch with the two-phase admission of FemBatch (prepare on worker threads)."""

    def __init__(self, capacity):
        super().__init__(capacity)
        self.prepared_on = set()

    def prepare(self, spec):
        import threading

        if getattr(spec, "untileable", False):
            raise ValueError("node has too many element corners")
        self.prepared_on.add(threading.get_ident())
        return SimpleNamespace(spec=spec, footprint=spec.bytes)

    def add_prepared(self, prep):
        return self.add(prep.spec)


def test_prepared_admission_keeps_fifo_and_isolates_untileable_jobs():
    b = FakePrepBatch(capacity=100)
    srv = SimServer(batch=b, frame=60, prep_workers=3)
    jobs = [Job(i, spec(i, h, 40), 0.0) for i, h in enumerate([30, 200, 59, 60, 500])]
    bad = spec(9, 60, 10)
    bad.untileable = True
    jobs.append(Job(9, bad, 0.0))
    res = {r.job_id: r for r in srv.run(jobs)}
    assert sorted(res) == [0, 1, 2, 3, 4, 9]
    assert not res[9].ok and all(res[i].ok for i in range(5))
    assert b.admit_order == [0, 1, 2, 3, 4]
    assert b.max_resident == 2
    assert len(b.prepared_on) >= 1


def test_frames_shorten_near_a_horizon():
    """A job is retired at most min_frame - 1 substeps past its horizon: the frame
    halves (60, 30, 15, 8) when the closest job needs fewer substeps."""
    b = FakeBatch(capacity=100)
    srv = SimServer(batch=b, frame=60, min_frame=8)
    assert srv.frame_lengths == [60, 30, 15, 8]
    res = srv.run([Job(0, spec(0, 100, 40), 0.0)])
    assert res[0].ok
    # warm-up + 100 substeps = 101: 60 + 30 + 8 + 8 (overshoot 5 < 8)
    assert b.frames == [60, 30, 8, 8]
    assert sum(b.frames) - 101 < 8


def test_default_frames_and_prep_workers(monkeypatch):
    """By default frames halve down to 15 substeps (60, 30, 15), so a job waits at most
    14 substeps past its horizon; preparation takes half the CPU share the process may
    run on, at most 8 threads and at least one, and OMP_NUM_THREADS caps that share."""
    srv = SimServer(batch=FakeBatch(capacity=100), frame=60)
    assert srv.frame_lengths == [60, 30, 15]
    assert SimServer(batch=FakeBatch(capacity=100), frame=10).frame_lengths == [10]
    for share, omp, want in ((16, "", 8), (32, "", 8), (6, "", 3), (1, "", 1), (16, "4", 2), (16, "x", 8)):
        monkeypatch.setattr(server_mod.os, "sched_getaffinity", lambda pid, n=share: set(range(n)))
        monkeypatch.setenv("OMP_NUM_THREADS", omp)
        assert default_prep_workers() == want, (share, omp)


class FakeAsyncBatch(FakePrepBatch):
    """FakePrepBatch with asynchronous retirement: a retired job's result exists only
    once the frames queued before its retirement have run (wait_launches), and the
    arrays handed out are filled by retire_wait -- as FemBatch.retire does."""

    def __init__(self, capacity):
        super().__init__(capacity)
        self.arena = SimpleNamespace(track_busy=lambda on=True: None, busy=lambda: (0.0, 0))
        self.queued = 0  # frames launched and not yet waited for
        self.max_queued = 0
        self.pending = []

    def step(self, n, graph=True, frame=None):
        super().step(n, graph, frame)
        self.queued += 1
        self.max_queued = max(self.max_queued, self.queued)

    def wait_launches(self, keep=0):
        self.queued = min(self.queued, keep)

    def retire(self, sid, velocities=True):
        s = self.scenes[sid]
        x = np.empty((4, 2), np.float32)
        v = np.empty((4, 2), np.float32) if velocities else None
        self.pending.append((x, v, s["spec"], self.queued))
        self.remove(sid)
        return x, v

    def retire_wait(self):
        for x, v, sp, q in self.pending:
            assert self.queued <= 1, "the frame a retirement follows must be done first"
            x[:] = np.nan if sp.poison else float(sp.name)
            if v is not None:
                v[:] = 0.0
        self.pending.clear()


def test_overlapped_retirement_keeps_results_and_two_frames_in_flight():
    b = FakeAsyncBatch(capacity=100)
    srv = SimServer(batch=b, frame=60, prep_workers=2)
    jobs = [Job(i, spec(i, h, 30), 0.0) for i, h in enumerate([30, 200, 59, 60, 500, 90])]
    bad = spec(7, 60, 10, poison=True)
    jobs.append(Job(7, bad, 0.0))
    res = {r.job_id: r for r in srv.run(jobs)}
    assert sorted(res) == [0, 1, 2, 3, 4, 5, 7]
    assert all(res[i].ok for i in range(6)) and not res[7].ok
    for i in range(6):
        assert np.all(res[i].positions == float(i))
    assert b.max_queued == 2  # the next frame is queued before the previous one is waited for
    assert b.admit_order[:3] == [0, 1, 2]
