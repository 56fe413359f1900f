"""Served form: a request loop that continuously batches simulation jobs.

The reference's ExplicitLoop drives one fixed set of domains from construction
to exit (flatworld/explicitloop.py:71-589); nothing can join or leave.  Here a
job -- one FEM body, its ground, its material and a horizon in substeps --
arrives at any time and is:

1. queued in arrival order;
2. admitted once its HBM footprint fits in the arena (head-of-line jobs
   wait; nothing is evicted) -- the host does this while the current frame
   runs on the device, and the job joins at the next frame;
3. advanced together with every other resident job, one fused launch per
   substep, frames replayed as a hipGraph (its zero-dt warm-up is simply its
   first substep in the arena);
4. retired when it reaches its horizon: its final state is copied out and its
   memory is released (stream-ordered, without stalling the running frame).
   With a batch that can retire asynchronously, the copies of the jobs a frame
   finished are queued behind that frame, the next frame is queued behind them,
   and only then does the host wait -- for the copies, not for the next frame --
   so the device does not idle while results come back.
   A frame is shortened (``frame``, then halves down to ``min_frame``
   substeps) when the resident job closest to its horizon needs fewer
   substeps, so a job waits at most ``min_frame - 1`` substeps past its horizon
   for retirement; each frame length is a cached graph of its own.

GPU-busy time is measured on the device: HIP events bracket every frame on the
arena stream (``Arena.track_busy``), so host admission and retirement between
frames do not count as busy.

Per-job latency is arrival -> result copied to the host.  A job's result is
bit-identical whether it runs alone or among others, because no device sum
depends on the resident set.  A job whose state turns non-finite (for
example a Neo-Hookean element inverting: log of a negative J) is reported as
failed at retirement without touching the other jobs.
"""

import collections
import os
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .batch import FemBatch


def default_prep_workers():
    """Preparation threads for this process's CPU share: half the cores it may run on
    (at most 8).  The request loop, the HIP runtime's threads and the page-locked
    copies need the other half; with every core preparing, the loop thread is
    descheduled between frames and the device idles (on a 16-core share, 16
    workers gave 970 jobs/s on a 300-job burst, 4 to 8 workers 1500 to 1800)."""
    try:
        share = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        share = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        share = min(share, int(omp))
    return max(1, min(8, share // 2))


@dataclass
class Job:
    job_id: int
    spec: object  # SceneSpec with horizon > 0
    arrival: float = 0.0  # seconds after the server clock starts


@dataclass
class JobResult:
    job_id: int
    arrival: float
    admitted: float
    finished: float
    horizon: int
    n_elems: int
    ok: bool
    positions: np.ndarray = field(repr=False, default=None)
    velocities: np.ndarray = field(repr=False, default=None)
    released: float = None  # when its arena memory was freed (at or before `finished`)

    @property
    def latency(self):
        return self.finished - self.arrival


class SimServer:
    """Continuous-batching request loop over one FemBatch.

    ``prep_workers`` host threads tile and pack arriving jobs (the batch's
    ``prepare``) in the background; admission then only copies the packed block
    into the arena.  A batch without ``prepare`` is admitted with ``add``.
    """

    def __init__(self, batch=None, frame=60, keep_results=True, clock=time.perf_counter, prep_workers=None,
                 min_frame=15, **batch_kw):
        self.batch = batch or FemBatch(frame=frame, **batch_kw)
        self.frame = int(frame)
        min_frame = max(1, min(int(min_frame), self.frame))
        self.frame_lengths = [self.frame]
        while self.frame_lengths[-1] > min_frame:
            self.frame_lengths.append(max(min_frame, self.frame_lengths[-1] // 2))
        self.keep_results = keep_results
        self.clock = clock
        if prep_workers is None:
            prep_workers = default_prep_workers()
        self.prep_workers = prep_workers if hasattr(self.batch, "prepare") else 0
        self.reset_stats()

    def reset_stats(self):
        self.results = []
        self.frames = 0
        self.substeps = 0
        self.busy_time = 0.0
        self.elem_substeps = 0

    def _progress(self, sid, job):
        """(substeps done after the warm-up, finished) of a resident job, from the
        substeps launched since its admission (the first is its warm-up), as the
        arena counts them."""
        ran = self._launched.get(sid, 0)
        done = min(max(ran - 1, 0), job.spec.horizon)
        return done, ran - 1 >= job.spec.horizon

    def _next_frame(self, resident):
        """The longest frame length that does not carry the job closest to its horizon past it."""
        rem = None
        for sid, (job, _) in resident.items():
            done = self._progress(sid, job)[0]
            r = job.spec.horizon - done + (1 if done == 0 else 0)  # + its warm-up substep
            rem = r if rem is None else min(rem, r)
        for n in self.frame_lengths:
            if n <= rem:
                return n
        return self.frame_lengths[-1]

    def run(self, jobs, max_wall=None):
        """Serve ``jobs`` (arrival times relative to the start of this call); returns the results.

        Up to two frames are queued on the device.  While they run, the host
        admits the jobs that arrived meanwhile (tiling and packing on the
        preparation threads, uploads queued behind the frames on the arena
        stream), so admission work overlaps device work; the admitted jobs join at
        the next frame.
        """
        pending = collections.deque(sorted(jobs, key=lambda j: j.arrival))
        waiting = collections.deque()
        resident = {}  # scene id -> (job, admitted time)
        t0 = self.clock()
        b = self.batch
        for job in pending:
            if job.spec.horizon <= 0:
                raise ValueError(f"job {job.job_id}: served jobs need a positive horizon")

        self._launched = {}  # scene id -> substeps launched since admission
        pool = ThreadPoolExecutor(self.prep_workers) if self.prep_workers > 0 else None
        prepared = {}  # job id -> Future of the job's prepared scene

        def admit():
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                job = pending.popleft()
                waiting.append(job)
                if pool is not None:
                    prepared[job.job_id] = pool.submit(b.prepare, job.spec)
            free = b.free_bytes()
            while waiting:
                job = waiting[0]
                fut = prepared.get(job.job_id)
                if fut is not None and not fut.done() and resident:
                    break  # head of the line still being tiled; admit it at a later call
                try:
                    prep = fut.result() if fut is not None else None
                except (RuntimeError, ValueError):
                    # the job's mesh cannot be tiled (e.g. a node with too many elements): it fails alone
                    waiting.popleft()
                    prepared.pop(job.job_id, None)
                    t = self.clock() - t0
                    self.results.append(JobResult(job.job_id, job.arrival, t, t, job.spec.horizon, job.spec.n_elems, False, released=t))
                    continue
                need = prep.footprint if prep is not None else b.footprint(job.spec)
                if need > free:
                    if not resident:
                        raise RuntimeError(f"job {job.job_id} needs {need} bytes; the empty arena has {free}")
                    break
                try:
                    sid = b.add_prepared(prep) if prep is not None else b.add(job.spec)
                except RuntimeError:
                    if resident:
                        break  # fragmented: wait for retirements
                    # the empty arena cannot take it: the job fails alone instead of blocking the line
                    waiting.popleft()
                    prepared.pop(job.job_id, None)
                    t = self.clock() - t0
                    self.results.append(JobResult(job.job_id, job.arrival, t, t, job.spec.horizon, job.spec.n_elems, False, released=t))
                    continue
                waiting.popleft()
                prepared.pop(job.job_id, None)
                resident[sid] = (job, self.clock() - t0)
                self._launched[sid] = 0
                free = b.free_bytes()

        def retire_queue():
            # jobs the launched frames finish: their readbacks queued behind those frames
            out = []
            for sid in list(resident.keys()):
                if not self._progress(sid, resident[sid][0])[1]:
                    continue
                job, adm = resident.pop(sid)
                self._launched.pop(sid, None)
                x, v = b.retire(sid, self.keep_results)
                out.append((job, adm, x, v, self.clock() - t0))
            return out

        def retire_finish(done):
            for job, adm, x, v, rel in done:
                ok = bool(np.isfinite(x).all())
                self.elem_substeps += job.spec.n_elems * job.spec.horizon
                self.results.append(
                    JobResult(job.job_id, job.arrival, adm, self.clock() - t0, job.spec.horizon, job.spec.n_elems, ok,
                              x if self.keep_results else None, v, rel)
                )

        self._retire_queue, self._retire_finish = retire_queue, retire_finish

        def retire():
            for sid in list(resident.keys()):
                steps, fin = self._progress(sid, resident[sid][0])
                if not fin:
                    continue
                job, adm = resident.pop(sid)
                self._launched.pop(sid, None)
                x = b.positions(sid)
                v = b.velocities(sid) if self.keep_results else None
                ok = bool(np.isfinite(x).all())
                b.remove(sid)
                self.elem_substeps += job.spec.n_elems * job.spec.horizon
                t = self.clock() - t0
                self.results.append(
                    JobResult(job.job_id, job.arrival, adm, t, job.spec.horizon, job.spec.n_elems, ok,
                              x if self.keep_results else None, v, t)
                )

        self._t0 = t0
        arena = getattr(b, "arena", None)
        busy0 = None
        if arena is not None:
            arena.track_busy(True)
            busy0 = arena.busy()[0]
        try:
            self._loop(pending, waiting, resident, admit, retire, max_wall)
        finally:
            if pool is not None:
                pool.shutdown(wait=True)
        self.wall = self.clock() - t0
        if busy0 is not None:
            self.busy_time = (arena.busy()[0] - busy0) * 1e-3  # GPU time of the frames, from HIP events
        return self.results

    def _launch(self, resident):
        n = self._next_frame(resident)
        self.batch.step(n, graph=True, frame=n)
        for sid in resident:
            self._launched[sid] += n
        self.frames += 1
        self.substeps += n

    def _loop(self, pending, waiting, resident, admit, retire, max_wall=None):
        b = self.batch
        t0 = self._t0
        in_flight = False
        overlap = hasattr(b, "retire") and hasattr(b, "wait_launches") and getattr(b, "arena", None) is not None
        while pending or waiting or resident:
            if max_wall is not None and self.clock() - t0 > max_wall:
                break
            if in_flight and overlap:
                # a frame is queued: queue the readbacks of the jobs it finishes, admit,
                # queue the next frame, then wait for the readbacks only
                done = self._retire_queue()
                admit()
                in_flight = bool(resident)
                if in_flight:
                    self._launch(resident)
                b.wait_launches(1 if in_flight else 0)  # the previous frame (two frames in flight at most)
                b.retire_wait()
                self._retire_finish(done)
                continue
            if in_flight:
                admit()  # overlaps the running frame
                b.sync()
                in_flight = False
                retire()
            admit()
            if resident:
                self._launch(resident)
                in_flight = True
            elif waiting:
                # nothing resident and the head of the line is still being tiled
                time.sleep(0.0002)
            elif pending:
                time.sleep(max(0.0, min(pending[0].arrival - (self.clock() - t0), 0.01)))
        if in_flight:
            b.sync()

    def stats(self):
        lat = np.array([r.latency for r in self.results]) if self.results else np.zeros(1)
        return {
            "jobs": len(self.results),
            "failed": sum(1 for r in self.results if not r.ok),
            "wall_s": self.wall,
            "frames": self.frames,
            "substeps": self.substeps,
            "gpu_busy_frac": self.busy_time / max(self.wall, 1e-9),
            "elem_substeps_per_s": self.elem_substeps / max(self.wall, 1e-9),
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_mean_s": float(lat.mean()),
        }


def poisson_arrivals(n, rate, rng):
    """n arrival times of a Poisson process with ``rate`` jobs per second."""
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
