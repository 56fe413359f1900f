"""Batched form: many independent one-body loops advanced by one launch per substep.

The reference batches only the domains of one ExplicitLoop (and silently drops
the ground contact of a loop holding more than one height field or voxel grid,
flatworld/mixedcontact.py:2140-2261), so thousands of scenes on their own
grounds need one loop each -- one set of launches per scene per substep.
``FemBatch`` keeps every scene resident in one HIP arena and advances all of
them with a single fused launch per substep; 60-substep frames replay as a
captured hipGraph.  Each scene keeps its own dt, damping, step counter,
activation-gate schedule and horizon, and its result does not depend on what
else is resident (no float atomics; every sum has a fixed order).
"""

from . import _native


class FemBatch:
    def __init__(self, device=0, capacity_bytes=0, tile_nodes=512, tile_slots=1024, block=256, grid_per_cu=0,
                 frame=60, track_disp=True):
        self.arena = _native.Arena(device=device, capacity_bytes=capacity_bytes, tile_nodes=tile_nodes,
                                   tile_slots=tile_slots, block=block, grid_per_cu=grid_per_cu)
        self.frame = int(frame)
        self.track_disp = track_disp
        self.specs = {}

    # ------------------------------------------------------------ scenes
    def footprint(self, spec):
        """Upper bound of the arena bytes ``spec`` needs (admission control)."""
        return self.arena.footprint(spec, self.track_disp)

    def free_bytes(self):
        return self.arena.free_bytes()

    def add(self, spec):
        """Admit a scene; its zero-dt warm-up runs in the next substep."""
        sid = self.arena.add(spec, track_disp=self.track_disp)
        self.specs[sid] = spec
        return sid

    def prepare(self, spec):
        """Host half of add(): tiling and packing, thread-safe, no GPU work; the packed
        image is then moved into page-locked memory on the same thread, so that
        admission only queues its upload."""
        return self.arena.prepare(spec, track_disp=self.track_disp).stage()

    def add_prepared(self, prep):
        """Device half of add(); raises RuntimeError (arena full) without consuming ``prep``."""
        sid = self.arena.add_prepared(prep)
        self.specs[sid] = prep.spec
        return sid

    def remove(self, sid):
        self.arena.remove(sid)
        self.specs.pop(sid, None)

    def retire(self, sid, velocities=True):
        """Remove a finished scene; its final (positions, velocities) arrays are filled
        by the next retire_wait(), which waits only for the launches before this call."""
        out = self.arena.retire(sid, velocities)
        self.specs.pop(sid, None)
        return out

    def retire_wait(self):
        self.arena.retire_wait()

    def wait_launches(self, keep=0):
        self.arena.wait_launches(keep)

    def progress(self, sid):
        """(substeps completed after the warm-up, reached horizon)."""
        return self.arena.progress(sid)

    # ------------------------------------------------------------ stepping
    def step(self, n, graph=True, diag=False, frame=None):
        """Advance every resident scene by n substeps (asynchronous): whole frames of
        ``frame`` substeps (default: the batch's) as graph replays, the rest as plain
        launches."""
        flags = _native.FW_STEP_DIAG if diag else 0
        frame = self.frame if frame is None else int(frame)
        if graph and frame > 0:
            full, rest = divmod(int(n), frame)
            for _ in range(full):
                self.arena.step_graph(frame, flags)
            if rest:
                self.arena.step(rest, -1.0, flags)
        else:
            self.arena.step(int(n), -1.0, flags)

    def step_per_scene(self, n):
        """The unbatched launch pattern: one launch per scene per substep (same results)."""
        self.arena.step(int(n), -1.0, _native.FW_STEP_PER_SCENE)

    def sync(self):
        self.arena.sync()

    # ------------------------------------------------------------ results
    def positions(self, sid):
        return self.arena.read(sid, _native.FIELD_X)

    def velocities(self, sid):
        return self.arena.read(sid, _native.FIELD_V)

    def read(self, sid, field):
        return self.arena.read(sid, field)

    def info(self):
        return self.arena.info()

    def close(self):
        self.arena.close()
