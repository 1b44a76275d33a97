"""Device-resident serving loop for covariance-style workloads (continuous batching).

Requests (one catalogue each, any grid / order / particle count) queue in arrival order.  Each
slot owns one device workspace (arena) that holds every intermediate buffer of the catalogue it
serves (assignment grids, half spectrum, sort / FFT / shell-map / GEMM scratch); arenas are kept
between requests and grow only when a request needs more, so a stream of requests does not
allocate and free tens of GB per catalogue (with a caching allocator that fragments, and frees
synchronise the device).  A request is admitted when a slot is free and its arena is large
enough, or can grow within the free device memory (hipMemGetInfo plus what torch's caching
allocator holds unused, plus the slot's own arena, released first); each admitted request runs
its whole pipeline on its slot's own HIP stream, so requests that arrive while others are on the
GPU start as soon as a slot and memory free up and their kernels overlap on the device.  Plans (rocFFT plans, exact triangle counts, bin and shell
tables) are built once per request shape and stay resident.  Completion is detected with HIP
events; results come back through pinned host buffers.  Latency = completion - arrival.

No counterpart in the reference, which measures one catalogue per process run.
"""
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from .pipeline import CataloguePlan, Measurement


ADMIT, DEFER, REJECT, WAIT = "admit", "defer", "reject", "wait"


def admission(now, arrival, free_slot, need_bytes, free_bytes, in_flight):
    """Decision for the request at the head of the queue.
    WAIT: it has not arrived yet; DEFER: no free slot, or not enough free device memory while
    other requests are in flight (they will release memory); REJECT: it cannot fit even on an
    idle device; ADMIT otherwise."""
    if arrival > now:
        return WAIT
    if not free_slot:
        return DEFER
    if need_bytes > free_bytes:
        return DEFER if in_flight > 0 else REJECT
    return ADMIT


@dataclass
class Request:
    rid: int
    measurement: Measurement
    pos: torch.Tensor              # (3, npart) float32 box units, on the device
    vel: torch.Tensor = None       # (3, npart) float32 km/s, on the device (redshift space)
    arrival: float = 0.0           # seconds, engine clock
    admitted: float = None
    done: float = None
    result: dict = None

    @property
    def npart(self):
        return int(self.pos.shape[1])

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


@dataclass
class _Slot:
    stream: torch.cuda.Stream
    req: Request = None
    event: torch.cuda.Event = None
    outputs: dict = None
    host: dict = field(default_factory=dict)
    arena: torch.Tensor = None

    @property
    def capacity(self):
        return 0 if self.arena is None else int(self.arena.numel())


class ServingEngine:
    def __init__(self, ctx, slots=4, reserve_bytes=2 << 30, free_bytes_fn=None):
        self.ctx = ctx
        self.slots = [_Slot(stream=torch.cuda.Stream(device=ctx.device)) for _ in range(slots)]
        self.queue = deque()
        self.finished = []
        self.plans = {}
        self.reserve = int(reserve_bytes)
        self._free_fn = free_bytes_fn          # None: the device's (no bound-method cycle)
        self.t0 = time.perf_counter()
        self.deferred = 0

    # ---------------------------------------------------------------- memory / plans
    def free_bytes(self):
        if self._free_fn is not None:
            return self._free_fn()
        free, _ = torch.cuda.mem_get_info(self.ctx.device)
        stats = torch.cuda.memory_stats(self.ctx.device)
        cached = stats.get("reserved_bytes.all.current", 0) - stats.get("allocated_bytes.all.current", 0)
        return free + max(0, cached)

    def release(self):
        """Drop the idle slots' workspaces (they return to torch's caching allocator)."""
        for s in self.slots:
            if s.req is None:
                s.arena = None

    def plan(self, m):
        p = self.plans.get(m)
        if p is None:
            p = self.plans[m] = CataloguePlan(self.ctx, m)
        return p

    def now(self):
        return time.perf_counter() - self.t0

    # ---------------------------------------------------------------- API
    def submit(self, req):
        if req.arrival is None or req.arrival == 0.0:
            req.arrival = self.now()
        self.queue.append(req)

    def in_flight(self):
        return sum(1 for s in self.slots if s.req is not None)

    def poll(self):
        """Retire finished requests, then admit queued ones (FIFO) while slots and memory
        allow.  Returns the requests finished by this call."""
        done = []
        for s in self.slots:
            if s.req is not None and s.event.query():
                r = s.req
                r.done = self.now()
                r.result = {f: (s.host[f][0].numpy().copy(), s.host[f][1].numpy().copy())
                            for f in s.host}
                s.req, s.outputs, s.host, s.event = None, None, {}, None
                done.append(r)
        self.finished.extend(done)
        while self.queue:
            r = self.queue[0]
            plan = self.plan(r.measurement)
            size = plan.arena_layout(r.npart)[3]
            free_slots = [s for s in self.slots if s.req is None]
            # a free slot whose arena already fits (the smallest such), else the free slot with
            # the largest arena, which grows
            fits = [s for s in free_slots if s.capacity >= size]
            slot = (min(fits, key=lambda s: s.capacity) if fits else
                    max(free_slots, key=lambda s: s.capacity) if free_slots else None)
            grow = slot is not None and slot.capacity < size
            need = (size + self.reserve) if grow else 0
            free = (self.free_bytes() + slot.capacity) if grow else 0
            d = admission(self.now(), r.arrival, slot is not None, need, free, self.in_flight())
            if d == REJECT:
                raise MemoryError(f"request {r.rid} needs {need} bytes, more than the device has")
            if d == ADMIT and grow and not self._grow(slot, size):
                d = DEFER if self.in_flight() > 0 else REJECT
                if d == REJECT:
                    raise MemoryError(f"request {r.rid}: a {size}-byte workspace does not fit")
            if d != ADMIT:
                self.deferred += d == DEFER and slot is not None
                break
            self.queue.popleft()
            self._launch(slot, r, plan)
        return done

    def _grow(self, slot, size):
        """Replace the (idle) slot's arena by one of `size` bytes; False if the device cannot
        hold it now."""
        slot.arena = None
        try:
            slot.arena = torch.empty(size, dtype=torch.uint8, device=self.ctx.device)
        except torch.OutOfMemoryError:
            torch.cuda.empty_cache()
            try:
                slot.arena = torch.empty(size, dtype=torch.uint8, device=self.ctx.device)
            except torch.OutOfMemoryError:
                return False
        return True

    def _launch(self, slot, r, plan):
        r.admitted = self.now()
        slot.req = r
        # the request's inputs are produced on the default stream
        slot.stream.wait_stream(torch.cuda.current_stream(self.ctx.device))
        r.pos.record_stream(slot.stream)
        if r.vel is not None:
            r.vel.record_stream(slot.stream)
        slot.outputs = plan.run(r.pos, r.vel, r.npart, stream=slot.stream, arena=slot.arena)
        with torch.cuda.stream(slot.stream):
            for f, (pout, bout) in slot.outputs.items():
                hp = torch.empty(pout.shape, dtype=pout.dtype, pin_memory=True)
                hb = torch.empty(bout.shape, dtype=bout.dtype, pin_memory=True)
                hp.copy_(pout, non_blocking=True)
                hb.copy_(bout, non_blocking=True)
                slot.host[f] = (hp, hb)
            slot.event = torch.cuda.Event()
            slot.event.record(slot.stream)

    def run(self, requests=None, poll_sleep=50e-6, progress=None):
        """Submit `requests` (their .arrival are offsets from now) and serve until all are
        finished.  Returns the finished requests in completion order.  progress: optional
        callable, given the number of this call's requests finished so far whenever some finish."""
        base = self.now()
        n0 = len(self.finished)
        for r in requests or []:
            r.arrival = base + (r.arrival or 0.0)
            self.queue.append(r)
        while self.queue or self.in_flight():
            if self.poll() and progress is not None:
                progress(len(self.finished) - n0)
            if self.queue or self.in_flight():
                time.sleep(poll_sleep)
        return self.finished[n0:]


def latency_summary(reqs):
    lat = np.array([r.latency for r in reqs if r.latency is not None]) * 1e3
    if lat.size == 0:
        return {}
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "n": int(lat.size)}
