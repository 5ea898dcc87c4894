"""Continuous-batching request loop for dbscan2d on one GPU.

No counterpart in the reference, which clusters one cloud per call.  Requests
(a cloud of float32 (N, 2) points with its own eps and min_samples) arrive at
any time; the server keeps a pooled HBM arena split into ``slots``, each able
to hold one batch (points, labels and the batched pipeline's workspace).
Whenever a slot is free, every queued request that fits is admitted into a
batch by its HBM footprint (first come, first served), copied into the slot
and clustered by one batched launch (dbscan2d_batched) on the slot's own
stream, so up to ``slots`` batches are on the device at once: requests that
arrive while others run are taken into the next batch as soon as a slot
frees, and their copies and kernels overlap the batches in flight.  A
request's labels are exactly what ``dbscan2d`` gives its cloud alone.

Admission is size-aware when ``big_slots`` > 0: requests of at least
``big_threshold`` points wait in their own queue and go only to the big
slots (each of ``big_points`` points), so the host-to-device copies and
kernels of multi-million-point requests never sit in front of small ones; a
big slot with no big request waiting takes small requests.

Typical use::

    server = Server(slots=3, slot_points=8 << 20, big_slots=3, big_points=24 << 20)
    server.submit(Request(0, X_host, eps=0.5, min_samples=10))
    server.drain()          # or call server.step() from your own loop
    labels = server.completed[0].labels
"""

from __future__ import annotations

import collections
import dataclasses
import time
from typing import Callable, Deque, Dict, List, Optional, Sequence

import torch

from . import _C

_ALIGN = 256


def _align(n: int) -> int:
    return (n + _ALIGN - 1) // _ALIGN * _ALIGN


@dataclasses.dataclass
class Request:
    """One dbscan2d call.  ``X`` is a float32 (N, 2) tensor, on the host
    (pinned memory lets its copy overlap) or on the server's device.  Its
    ``labels`` end up where ``X`` is, as ``dbscan2d`` returns them: in host
    memory for a host ``X``, on the device for a device-resident one (nothing
    crosses PCIe then).  ``arrival`` is the time the request is due (clock
    seconds); latency is measured from it to the moment its labels are
    complete in that memory.  ``error`` is None, or the text of a failed
    device-side check of its batch (its labels are then INT32_MIN)."""

    rid: int
    X: torch.Tensor
    eps: float
    min_samples: int
    arrival: float = 0.0
    labels: Optional[torch.Tensor] = None
    admitted: Optional[float] = None
    done: Optional[float] = None
    error: Optional[str] = None

    @property
    def n(self) -> int:
        return int(self.X.shape[0])

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class _Slot:
    """One batch's share of the arena: points, labels, workspace and its
    stream.  A batch's labels leave the arena in at most two copies: one
    device-to-host copy into a pinned tensor of the batch's own (torch's
    caching host allocator) when it holds host requests, and one
    device-to-device copy into a device tensor of its own when it holds
    device-resident requests.  Each request receives its slice of one of
    them, so retiring a batch copies nothing on the host (a request's labels
    keep that batch's buffer alive)."""

    def __init__(self, arena: torch.Tensor, offset: int, points: int, clouds: int,
                 ws_bytes: int, device: torch.device, big: bool = False):
        self.points, self.clouds, self.big = points, clouds, big
        x_bytes, l_bytes = _align(points * 8), _align(points * 4)
        self.X = arena.narrow(0, offset, points * 8).view(torch.float32).view(points, 2)
        self.labels = arena.narrow(0, offset + x_bytes, points * 4).view(torch.int32)
        self.workspace = arena.narrow(0, offset + x_bytes + l_bytes, ws_bytes)
        self.stream = torch.cuda.Stream(device)  # default priority
        self.event: Optional[torch.cuda.Event] = None
        # the batch's device error word, copied back with its labels
        self.err = torch.zeros(1, dtype=torch.int32, pin_memory=True)
        self.batch: List[Request] = []
        self.offsets: List[int] = []
        self.host_labels: Optional[torch.Tensor] = None
        self.dev_labels: Optional[torch.Tensor] = None

    @staticmethod
    def nbytes(points: int, ws_bytes: int) -> int:
        return _align(points * 8) + _align(points * 4) + _align(ws_bytes)


class Server:
    """Pooled-arena, continuous-batching dbscan2d server on one GPU.

    Args:
        device: the ROCm device.
        slots: batches of small requests that may be on the device at once.
        slot_points: points one such batch holds (without big slots: the
            largest admissible request).
        slot_clouds: requests one batch holds.
        cap_bytes: the resident HBM budget; the arena (all slots) must fit.
        clock: time source for admission and completion stamps.
        big_slots: slots reserved for requests of at least ``big_threshold``
            points (0: one queue, every slot takes every request).
        big_points: points one big slot holds (the largest admissible
            request when there are big slots).
        big_threshold: the size from which a request is big.
    """

    def __init__(self, device="cuda", slots: int = 3, slot_points: int = 32 << 20,
                 slot_clouds: int = 4096, cap_bytes: float = 200e9,
                 clock: Callable[[], float] = time.perf_counter, big_slots: int = 0,
                 big_points: Optional[int] = None, big_threshold: int = 1 << 20):
        if slots < 1 or slot_points < 1 or slot_clouds < 1 or big_slots < 0:
            raise ValueError("Server: slots, slot_points and slot_clouds must be >= 1")
        self.device = torch.device(device)
        self.slot_points, self.slot_clouds, self.clock = slot_points, slot_clouds, clock
        self.big_points = (big_points or slot_points) if big_slots else 0
        self.big_threshold = big_threshold if big_slots else None
        if big_slots and self.big_points < big_threshold:
            raise ValueError("Server: big_points must be at least big_threshold")
        ws = _C.dbscan2d_batched_workspace_bytes(slot_points, slot_clouds)
        slot_bytes = _Slot.nbytes(slot_points, ws)
        big_ws = _C.dbscan2d_batched_workspace_bytes(self.big_points, slot_clouds) if big_slots else 0
        big_bytes = _Slot.nbytes(self.big_points, big_ws) if big_slots else 0
        self.arena_bytes = slots * slot_bytes + big_slots * big_bytes
        if self.arena_bytes > cap_bytes:
            raise ValueError(f"Server: arena of {self.arena_bytes / 1e9:.1f} GB exceeds the "
                             f"{cap_bytes / 1e9:.1f} GB resident cap")
        self.arena = torch.empty(self.arena_bytes, dtype=torch.uint8, device=self.device)
        self.slots = [_Slot(self.arena, i * slot_bytes, slot_points, slot_clouds, ws, self.device)
                      for i in range(slots)]
        self.slots += [_Slot(self.arena, slots * slot_bytes + i * big_bytes, self.big_points,
                             slot_clouds, big_ws, self.device, big=True)
                       for i in range(big_slots)]
        self.bytes_per_point = slot_bytes / slot_points
        self.queue: Deque[Request] = collections.deque()      # small (or every) request
        self.big_queue: Deque[Request] = collections.deque()  # big requests
        self.free: Deque[_Slot] = collections.deque(self.slots)
        self.inflight: Deque[_Slot] = collections.deque()
        self.completed: Dict[int, Request] = {}
        self.failed: List[int] = []  # ids of requests whose batch failed a device check
        self.batches: List[int] = []  # requests per launched batch

    def footprint_bytes(self, n: int) -> int:
        """Arena bytes a request of n points occupies (points, labels and its
        share of the workspace)."""
        return int(n * self.bytes_per_point)

    def submit(self, req: Request) -> None:
        X = req.X
        if X.dim() != 2 or X.size(1) != 2:
            raise RuntimeError("X must have shape (N, 2)")
        if X.dtype != torch.float32:
            raise RuntimeError("X must be float32")
        if not req.eps > 0.0:
            raise RuntimeError("eps must be > 0")
        if req.min_samples < 1:
            raise RuntimeError("min_samples must be >= 1")
        big = self.big_threshold is not None and (req.n >= self.big_threshold or
                                                  req.n > self.slot_points)
        limit = self.big_points if big else self.slot_points
        if req.n > limit:
            raise ValueError(f"Server: request {req.rid} has {req.n} points; a slot holds "
                             f"{limit}")
        if req.n == 0:
            req.labels = torch.empty(0, dtype=torch.int32, device=X.device)
            req.admitted = req.done = self.clock()
            self.completed[req.rid] = req
            return
        if big:
            self.big_queue.append(req)
        else:
            self.queue.append(req)

    def busy(self) -> bool:
        return bool(self.queue) or bool(self.big_queue) or bool(self.inflight)

    def _source(self, slot: _Slot) -> Optional[Deque[Request]]:
        """The queue a free slot takes requests from, or None."""
        if slot.big and self.big_queue:
            return self.big_queue
        return self.queue if self.queue else None

    def _launch(self, slot: _Slot, queue: Deque[Request]) -> None:
        batch, offs, pts = [], [0], 0
        while (queue and len(batch) < slot.clouds and pts + queue[0].n <= slot.points):
            r = queue.popleft()
            batch.append(r)
            pts += r.n
            offs.append(pts)
        if not batch:  # the head request does not fit this slot
            self.free.append(slot)
            return
        now = self.clock()
        for r in batch:
            r.admitted = now
        on_device = [r.X.is_cuda for r in batch]
        host = (torch.empty(pts, dtype=torch.int32, pin_memory=True)
                if not all(on_device) else None)
        dev = None
        # device-resident inputs may still be being written on the caller's
        # stream: the slot's stream waits for it, and their memory is kept
        # until the slot's copies have read it
        if any(on_device):
            slot.stream.wait_stream(torch.cuda.current_stream(self.device))
            for r, d in zip(batch, on_device):
                if d:
                    r.X.record_stream(slot.stream)
        with torch.cuda.stream(slot.stream):
            _C.copy_clouds(slot.X, [r.X for r in batch], offs[:-1])
            _C.dbscan2d_batched_out(slot.X[:pts], offs, [r.eps for r in batch],
                                    [r.min_samples for r in batch], slot.labels,
                                    slot.workspace, slot.err)
            if host is not None:
                host.copy_(slot.labels[:pts], non_blocking=True)
            if any(on_device):
                # allocated on the slot's stream, which writes it; it is handed
                # out only after the batch's event has completed
                dev = slot.labels[:pts].clone()
            slot.event = torch.cuda.Event()
            slot.event.record(slot.stream)
        slot.batch, slot.offsets, slot.host_labels, slot.dev_labels = batch, offs, host, dev
        self.inflight.append(slot)
        self.batches.append(len(batch))

    def _retire(self, slot: _Slot) -> None:
        now = self.clock()
        bits = int(slot.err[0])
        why = _C.device_error_message(bits) if bits else None
        for r, o in zip(slot.batch, slot.offsets):
            r.labels = (slot.dev_labels if r.X.is_cuda else slot.host_labels)[o:o + r.n]
            r.done = now
            r.error = why
            if why:
                self.failed.append(r.rid)
            self.completed[r.rid] = r
        slot.batch, slot.offsets, slot.event = [], [], None
        slot.host_labels = slot.dev_labels = None
        self.free.append(slot)

    def step(self) -> bool:
        """Retires finished batches and launches new ones; never blocks.
        Returns True if anything happened."""
        progressed = False
        for slot in list(self.inflight):
            if slot.event.query():
                self.inflight.remove(slot)
                self._retire(slot)
                progressed = True
        for slot in list(self.free):
            queue = self._source(slot)
            if queue is not None:
                self.free.remove(slot)
                self._launch(slot, queue)
                progressed = True
        return progressed

    def drain(self) -> None:
        while self.busy():
            if not self.step():
                time.sleep(20e-6)


def run_open_loop(server: Server, requests: Sequence[Request],
                  arrivals: Sequence[float]) -> List[Request]:
    """Serves ``requests`` as an open-loop stream: request i is submitted once
    ``arrivals[i]`` seconds have passed since the start (never earlier), and
    its latency runs from that scheduled time.  Returns the requests."""
    order = sorted(range(len(requests)), key=lambda i: arrivals[i])
    t0 = server.clock()
    k = 0
    while k < len(order) or server.busy():
        now = server.clock() - t0
        while k < len(order) and arrivals[order[k]] <= now:
            r = requests[order[k]]
            r.arrival = t0 + arrivals[order[k]]
            server.submit(r)
            k += 1
        if not server.step() and (k == len(order) or arrivals[order[k]] > now + 50e-6):
            time.sleep(20e-6)
    return list(requests)


def latency_summary(reqs: Sequence[Request]) -> dict:
    """p50 / p99 / max latency in ms and completed clouds per second."""
    lat = sorted(r.latency for r in reqs)
    if not lat:
        return {"requests": 0}

    def pct(p):
        return lat[min(len(lat) - 1, int(round(p / 100.0 * (len(lat) - 1))))] * 1e3

    span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
    return {"requests": len(lat), "p50_ms": round(pct(50), 3), "p99_ms": round(pct(99), 3),
            "max_ms": round(lat[-1] * 1e3, 3),
            "clouds_per_s": round(len(lat) / span, 1) if span > 0 else None}
