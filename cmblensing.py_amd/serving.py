"""Serving: continuous batching of LenseFlow requests on one GPU.

The reference serves one call at a time (lense_flow, lense_flow.py:289-292);
this module keeps that call's meaning per request -- a request is one T map or
one Q/U pair, lensed (direction +1), unlensed (-1) or adjoint-lensed by one of
a set of shared lensing potentials -- and runs many requests together:

* admission: a request is admitted when its input and output maps, plus the
  operators of its phi if they are not resident yet, fit the HBM budget
  (min of the configured budget and the device's free memory at start-up);
  admission is first come first served, and resident phi operators that no
  queued or running request uses are evicted (least recently used) to make
  room;
* resident operators: the phi derivatives and the p(t) table of a (phi, N)
  are prepared once, on first use, and shared by every later request for the
  same phi (the hit rate is reported);
* continuous batching: admitted requests wait in buckets of one (N, C,
  direction, adjoint); whenever a stream has room, the bucket whose oldest
  request arrived first is launched as one batch of up to the bucket's batch
  size, so requests that arrive while others run are packed into the next
  launch and batches of different buckets overlap on the device's streams;
  when a bucket holds more than one batch while several stream slots are
  free, its next batches are formed together and ordered by phi across all
  of them before the split, so requests sharing a phi fall into one batch;
  a batch's positions are ordered by phi, so those requests read its p(t)
  rows together;
* each batch is one replay of a hipGraph captured once per (stream, bucket,
  batch class), with the request maps gathered into and scattered out of a
  staging buffer by the graph's own kernels;
* latency is arrival -> completion, the completion time read from a HIP
  event recorded after the batch on its stream;
* Wiener-filter requests (``WFRequest``: data, phi and operators of one
  lensed Wiener filter, wiener_filter.py:10-26) pass the same admission and
  run one at a time in a lane of their own stream.  The solve is a process
  (``wiener_filter_process``) that yields an event wherever the PCG needs a
  host scalar; the scheduler advances it when that event has completed, so
  the host never blocks on a solve and LenseFlow batches keep being formed
  and overlap with it on the device.

The scheduler (``Server``) is independent of the device: ``GpuBackend`` runs
the batches through libcmbl.so, and any object with the same methods (tests
use a host-side one) can stand in for it.
"""
from __future__ import annotations

import math
import sys
import threading
import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np
import torch

from .constants import FORWARD_LENSE, INVERSE_LENSE

GiB = 1 << 30
DEFAULT_MAX_BATCH = {16: 256, 32: 256, 64: 256, 128: 128, 256: 64, 512: 32}


@dataclass(eq=False)
class Request:
    """One LenseFlow request: x is [C, N, N] (C = 1 for T, 2 for Q/U) on the
    serving device; phi_id names one of the server's shared potentials."""
    x: torch.Tensor
    phi_id: int
    direction: int = FORWARD_LENSE
    adjoint: bool = False
    arrival: Optional[float] = None  # time.perf_counter() seconds; set on submit if None
    rid: int = -1
    # filled in by the server
    out: Optional[torch.Tensor] = None
    done: Optional[float] = None
    state: str = "new"  # new, waiting, queued, running, done, cancelled, rejected
    error: Optional[str] = None
    nbytes: int = 0

    @property
    def N(self) -> int:
        return int(self.x.shape[-1])

    @property
    def C(self) -> int:
        return int(self.x.shape[0])

    @property
    def key(self):
        return (self.N, self.C, int(self.direction), bool(self.adjoint))

    @property
    def latency(self) -> Optional[float]:
        if self.done is None or self.arrival is None:
            return None
        return self.done - self.arrival


@dataclass(eq=False)
class WFRequest:
    """One lensed Wiener-filter solve (wiener_filter.py:10-26): ``data`` is an
    EB (or T) FOURIER field on the serving device (maps with a PixelMask),
    ``ops`` = (field_covariance, noise_covariance, mask, beam); the field dtype
    sets the operator precision (complex64: FP32 operator, FP64 CG).  The
    result is ``out`` (the filtered field) and ``info`` (CGInfo)."""
    data: object
    phi_id: int
    ops: tuple
    n: int = 7
    rtol: float = 1e-6
    maxiter: int = 500
    arrival: Optional[float] = None
    rid: int = -1
    out: object = None
    info: object = None
    done: Optional[float] = None
    state: str = "new"
    error: Optional[str] = None
    nbytes: int = 0

    @property
    def N(self) -> int:
        m = getattr(self.data, self.data._matrix_names()[0])
        return int(m.shape[-2])

    @property
    def C(self) -> int:
        return len(self.data._matrix_names())

    @property
    def latency(self) -> Optional[float]:
        if self.done is None or self.arrival is None:
            return None
        return self.done - self.arrival


@dataclass(eq=False)
class _WFToken:
    gen: object
    finished: bool = False
    result: object = None
    error: Optional[str] = None
    end: object = None
    thread: object = None


@dataclass(eq=False)
class _Batch:
    token: object
    requests: List[Request]
    stream: int
    key: tuple


@dataclass(eq=False)
class _PhiEntry:
    handle: object
    nbytes: int
    refs: int = 0


class Server:
    """Admission, resident phi operators and continuous batching over a
    backend (see the module docstring).  Drive it with submit() and step(),
    or with serve() for a timed arrival schedule."""

    def __init__(self, backend, phi_source: Callable[[int, int], torch.Tensor],
                 hbm_budget: int = 64 * GiB, max_inflight: int = 2, min_det: float = 0.0):
        self.backend = backend
        # phi whose magnification determinant det(I + t Hess phi) falls to
        # min_det or below somewhere in t in [0, 1] (a caustic: M(t) is
        # singular and the flow is not defined) are refused
        self.min_det = min_det
        self.bad_phi: Dict[tuple, float] = {}
        self.phi_source = phi_source
        free = backend.free_bytes()
        self.budget = int(min(hbm_budget, free)) if free is not None else int(hbm_budget)
        self.max_inflight = max_inflight
        self.used = backend.fixed_bytes()
        self.peak_used = self.used
        self.waiting: deque = deque()
        self.buckets: "OrderedDict[tuple, deque]" = OrderedDict()
        self.inflight: List[_Batch] = []
        self.per_stream = [0] * backend.nstreams
        self.phis: "OrderedDict[tuple, _PhiEntry]" = OrderedDict()
        self.hits = self.misses = self.evictions = 0
        self.completed: List[Request] = []
        self.completed_wf: List[WFRequest] = []
        self.wf_queue: deque = deque()
        self.wf_running = 0
        self.wf_slots = getattr(backend, "wf_slots", 1)
        self.batch_sizes: List[int] = []
        self._next_rid = 0
        self._by_rid: Dict[int, Request] = {}

    # -- requests ----------------------------------------------------------
    def submit(self, req) -> object:
        if isinstance(req, WFRequest):
            return self._submit_common(req)
        if req.N not in self.backend.sizes or req.C not in (1, 2):
            raise ValueError(f"unsupported request shape {tuple(req.x.shape)}")
        if req.direction not in (FORWARD_LENSE, INVERSE_LENSE):
            raise ValueError("direction must be FORWARD_LENSE (+1) or INVERSE_LENSE (-1)")
        return self._submit_common(req)

    def _submit_common(self, req):
        if req.arrival is None:
            req.arrival = time.perf_counter()
        if req.rid < 0:
            req.rid = self._next_rid
            self._next_rid += 1
        req.nbytes = (self.backend.wf_bytes(req) if isinstance(req, WFRequest)
                      else self.backend.request_bytes(req))
        req.state = "waiting"
        self._by_rid[req.rid] = req
        self.waiting.append(req)
        return req

    def cancel(self, rid: int) -> bool:
        """Cancel a request that has not completed.  A waiting or queued
        request is dropped at once; a running one is marked and its result
        discarded when its batch completes (a launched batch is not cut)."""
        req = self._by_rid.get(rid)
        if req is None or req.state in ("done", "cancelled", "rejected"):
            return False
        if req.state == "waiting":
            self.waiting.remove(req)
            req.state = "cancelled"
        elif req.state == "queued":
            if isinstance(req, WFRequest):
                self.wf_queue.remove(req)
            else:
                self.buckets[req.key].remove(req)
            self._release(req)
            req.state = "cancelled"
        else:  # running
            req.state = "cancelling"
        return True

    # -- admission -----------------------------------------------------------
    def _phi_cost(self, req) -> int:
        if isinstance(req, WFRequest):
            return 0  # prepared on the Wiener-filter lane's own plans (counted in wf_bytes)
        k = (req.phi_id, req.N)
        return 0 if k in self.phis else self.backend.phi_bytes(req.N)

    def _evict_for(self, need: int, keep) -> bool:
        """Evict unreferenced resident phi (LRU first, never `keep`) until
        `need` bytes fit."""
        while self.used + need > self.budget:
            victim = next((k for k, e in self.phis.items() if e.refs == 0 and k != keep), None)
            if victim is None:
                return False
            e = self.phis.pop(victim)
            self.backend.release_phi(e.handle)
            self.used -= e.nbytes
            self.evictions += 1
        return True

    def _admit(self):
        while self.waiting:
            req = self.waiting[0]
            need = req.nbytes + self._phi_cost(req)
            if need > self.budget - self.backend.fixed_bytes():
                self.waiting.popleft()
                req.state, req.error = "rejected", "request does not fit the HBM budget"
                continue
            if self.used + need > self.budget and not self._evict_for(need, (req.phi_id, req.N)):
                return  # wait for running batches to free memory
            self.waiting.popleft()
            kb = (req.phi_id, req.N)
            if kb in self.bad_phi:
                req.state = "rejected"
                req.error = f"phi {req.phi_id} has a caustic (min det {self.bad_phi[kb]:.3g})"
                continue
            if isinstance(req, WFRequest):
                self.used += req.nbytes
                self.peak_used = max(self.peak_used, self.used)
                req.state = "queued"
                self.wf_queue.append(req)
                continue
            k = (req.phi_id, req.N)
            ent = self.phis.get(k)
            if ent is None:
                self.misses += 1
                nb = self.backend.phi_bytes(req.N)
                handle = self.backend.prepare_phi(req.N, self.phi_source(req.phi_id, req.N))
                md = getattr(handle, "min_det", None)
                if md is not None and md <= self.min_det:
                    self.backend.release_phi(handle)
                    self.bad_phi[k] = md
                    req.state = "rejected"
                    req.error = f"phi {req.phi_id} has a caustic (min det {md:.3g})"
                    continue
                ent = _PhiEntry(handle, nb)
                self.phis[k] = ent
                self.used += nb
            else:
                self.hits += 1
                self.phis.move_to_end(k)
            ent.refs += 1
            req.out = self.backend.alloc_output(req)
            self.used += req.nbytes
            self.peak_used = max(self.peak_used, self.used)
            req.state = "queued"
            self.buckets.setdefault(req.key, deque()).append(req)

    def _release(self, req):
        self.used -= req.nbytes
        if not isinstance(req, WFRequest):
            self.phis[(req.phi_id, req.N)].refs -= 1

    # -- batching ------------------------------------------------------------
    def _launch(self) -> int:
        launched = 0
        while self.wf_queue and self.wf_running < self.wf_slots:
            req = self.wf_queue.popleft()
            req.state = "running"
            token = self.backend.launch_wf(req, self.phi_source(req.phi_id, req.N))
            self.inflight.append(_Batch(token, [req], -1, ("wf",)))
            self.wf_running += 1
            launched += 1
        while True:
            s = self._free_stream()
            if s is None:
                break
            key, q = None, None
            for k, dq in self.buckets.items():
                if dq and (q is None or dq[0].arrival < q[0].arrival):
                    key, q = k, dq
            if q is None:
                break
            cap = self.backend.max_batch(key)
            # batch positions ordered by phi: requests that share a phi then
            # read the same rows of its p(t) table at about the same time
            # (their row-pass workgroups lie a fixed stride apart, on one
            # XCD), so its L2 serves the repeats; a request's result does not
            # depend on its batch or position
            if len(q) <= cap:
                reqs = sorted((q.popleft() for _ in range(len(q))), key=lambda r: r.phi_id)
                launched += self._start(key, reqs, s)
                continue
            # a bucket holding more than one batch while several slots are
            # free: its next batches are formed together and launched now,
            # ordered by phi across all of them before the split, so requests
            # that share a phi fall into one batch; no request waits for the
            # grouping.  (The slot list is built only here: the loop runs on
            # every scheduler step, and allocating on each one set off garbage
            # collections that held the interpreter lock and stalled the
            # Wiener-filter lane's thread.)
            slots = self._free_slots()
            nbat = min(len(slots), -(-len(q) // cap))
            m = min(len(q), nbat * cap)
            reqs_all = sorted((q.popleft() for _ in range(m)), key=lambda r: r.phi_id)
            for i in range(nbat):
                launched += self._start(key, reqs_all[i * cap:(i + 1) * cap], slots[i])
        return launched

    def _start(self, key, reqs, s) -> int:
        handles = [self.phis[(r.phi_id, r.N)].handle for r in reqs]
        for r in reqs:
            r.state = "running"
        token = self.backend.launch(key, reqs, handles, s)
        self.inflight.append(_Batch(token, reqs, s, key))
        self.per_stream[s] += 1
        self.batch_sizes.append(len(reqs))
        return 1

    def _free_stream(self) -> Optional[int]:
        best = None
        for s, n in enumerate(self.per_stream):
            if n < self.max_inflight and (best is None or n < self.per_stream[best]):
                best = s
        return best

    def _free_slots(self):
        """Streams with room, in the order batches would be given to them
        one at a time (least loaded first, up to max_inflight each)."""
        load = list(self.per_stream)
        slots = []
        while True:
            best = None
            for s, n in enumerate(load):
                if n < self.max_inflight and (best is None or n < load[best]):
                    best = s
            if best is None:
                return slots
            slots.append(best)
            load[best] += 1

    def _poll(self) -> int:
        finished = 0
        still = []
        for b in self.inflight:
            if not self.backend.ready(b.token):
                still.append(b)
                continue
            t = self.backend.finish_time(b.token)
            if b.stream < 0:  # the Wiener-filter lane
                self.wf_running -= 1
                r = b.requests[0]
                self._release(r)
                err = self.backend.wf_result(b.token, r)
                if r.state == "cancelling":
                    r.state, r.out, r.info = "cancelled", None, None
                elif err is not None:
                    r.state, r.error = "failed", err
                else:
                    r.done, r.state = t, "done"
                    self.completed_wf.append(r)
                    finished += 1
                continue
            for r in b.requests:
                self._release(r)
                if r.state == "cancelling":
                    r.state, r.out = "cancelled", None
                    continue
                r.done, r.state = t, "done"
                self.completed.append(r)
                finished += 1
            self.per_stream[b.stream] -= 1
        self.inflight = still
        return finished

    def step(self) -> int:
        """One scheduling pass: retire finished batches, admit, launch.
        Returns the number of batches retired plus launched."""
        n = self._poll()
        self._admit()
        return n + self._launch()

    def busy(self) -> bool:
        return bool(self.waiting or self.inflight or self.wf_queue or any(self.buckets.values()))

    def drain(self):
        while self.busy():
            if not self.step():
                time.sleep(0)

    def serve(self, requests: List[Request], offsets: List[float]) -> dict:
        """Open-loop run: request i arrives `offsets[i]` seconds after the
        start (its arrival time is the scheduled one, so scheduling delay
        counts toward its latency).  Returns stats()."""
        order = np.argsort(np.asarray(offsets), kind="stable")
        t0 = self.backend.start_clock()
        i = 0
        n = len(requests)
        while i < n or self.busy():
            now = time.perf_counter()
            while i < n and t0 + offsets[order[i]] <= now:
                r = requests[order[i]]
                r.arrival = t0 + offsets[order[i]]
                self.submit(r)
                i += 1
            if not self.step():
                # an idle pass hands the interpreter lock to the Wiener-filter
                # lane's thread, which needs it between its launches
                time.sleep(0)
        return self.stats()

    # -- reporting -------------------------------------------------------------
    def stats(self) -> dict:
        lat = np.array([r.latency for r in self.completed], dtype=np.float64)
        out = {"completed": len(self.completed), "batches": len(self.batch_sizes),
               "mean_batch": float(np.mean(self.batch_sizes)) if self.batch_sizes else 0.0,
               "phi_hits": self.hits, "phi_misses": self.misses,
               "phi_hit_rate": self.hits / max(1, self.hits + self.misses),
               "phi_evictions": self.evictions,
               "hbm_budget_bytes": self.budget, "hbm_peak_bytes": self.peak_used}
        if len(lat):
            first = min(r.arrival for r in self.completed)
            last = max(r.done for r in self.completed)
            out.update({"p50_ms": float(np.percentile(lat, 50) * 1e3),
                        "p99_ms": float(np.percentile(lat, 99) * 1e3),
                        "max_ms": float(lat.max() * 1e3),
                        "maps_per_s": len(lat) / max(last - first, 1e-12)})
        if self.completed_wf:
            wl = np.array([r.latency for r in self.completed_wf], dtype=np.float64)
            out.update({"wf_completed": len(wl), "wf_p50_ms": float(np.percentile(wl, 50) * 1e3),
                        "wf_p99_ms": float(np.percentile(wl, 99) * 1e3),
                        "wf_iterations_mean": float(np.mean([max(r.info.iterations)
                                                             for r in self.completed_wf]))})
        return out


def _batch_class(count: int, cap: int) -> int:
    c = 1 << max(0, math.ceil(math.log2(max(1, count))))
    return min(c, cap)


class GpuBackend:
    """Runs batches through libcmbl.so graph executors on `nstreams` streams
    of default priority.  Per (stream, N) one plan (scratch and staging for
    the N's largest batch); per (stream, bucket, power-of-two batch class) one
    executor, created on first use or by warmup()."""

    def __init__(self, device="cuda:0", dtype=torch.float32, theta_pix: float = 2.0,
                 nsteps: int = 7, nstreams: int = 2, max_batch: Optional[Dict[int, int]] = None,
                 sizes=(64, 128, 256, 512)):
        from . import _lib
        self._lib = _lib
        _lib.load()
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.dtype, self.theta_pix, self.nsteps = dtype, float(theta_pix), nsteps
        self.sizes = tuple(sizes)
        self.mb = {N: (max_batch or DEFAULT_MAX_BATCH).get(N, 32) for N in self.sizes}
        self.nstreams = nstreams
        self.streams = [torch.cuda.Stream(self.device) for _ in range(nstreams)]
        self.prep_stream = torch.cuda.Stream(self.device)
        self.s = 4 if dtype == torch.float32 else 8
        self.plans: dict = {}
        self.execs: dict = {}
        self._fixed = 0
        self._anchor_ev = None
        self._anchor_t = 0.0
        self._wf_stream = None
        self._switch_interval = None    # CPython's switch interval while a lane thread runs
        self._warm = False              # every default bucket's executors captured
        self.wf_slots = 1

    # memory model ---------------------------------------------------------------
    def free_bytes(self) -> Optional[int]:
        free, _ = torch.cuda.mem_get_info(self.device)
        return int(free * 0.9)

    def fixed_bytes(self) -> int:
        """Plan scratch + staging of every (stream, N) plan (4 map stacks of
        the plan's max_maps), counted up front."""
        return sum(4 * 2 * self.mb[N] * N * N * self.s for N in self.sizes) * self.nstreams

    def request_bytes(self, req: Request) -> int:
        return 2 * req.C * req.N * req.N * self.s

    def phi_bytes(self, N: int) -> int:
        return 5 * N * N * 8 + (2 * self.nsteps + 1) * 2 * N * N * self.s

    def max_batch(self, key) -> int:
        return self.mb[key[0]]

    # resources --------------------------------------------------------------------
    def _plan(self, si: int, N: int):
        p = self.plans.get((si, N))
        if p is None:
            p = self._lib.Plan(N, self.theta_pix, self.dtype, 2 * self.mb[N], self.device)
            self.plans[(si, N)] = p
        return p

    def _exec(self, si: int, key, cls: int):
        ek = (si, key, cls)
        ex = self.execs.get(ek)
        if ex is None:
            # a graph capture must not overlap the Wiener-filter lane's device
            # work (its synchronisations would invalidate the capture): wait
            # for a running lane first (only for a bucket warmup() did not
            # cover -- launch_wf warms the default buckets before the lane's
            # first solve)
            for t in getattr(self, "_wf_threads", []):
                t.join()
            N, C, direction, adjoint = key
            ex = self._lib.Executor(self._plan(si, N), cls, C, self.nsteps, direction, adjoint)
            self.execs[ek] = ex
        return ex

    def warmup(self, keys=None):
        """Create every executor (all batch classes) for `keys` (default: every
        bucket of the served sizes), so no graph is captured while serving."""
        default = keys is None
        if keys is None:
            keys = [(N, C, d, False) for N in self.sizes for C in (1, 2)
                    for d in (FORWARD_LENSE, INVERSE_LENSE)]
        for key in keys:
            cap = self.mb[key[0]]
            classes = sorted({_batch_class(c, cap) for c in range(1, cap + 1)})
            for si in range(self.nstreams):
                for cls in classes:
                    self._exec(si, key, cls)
        torch.cuda.synchronize(self.device)
        self._warm = self._warm or default

    def prepare_phi(self, N: int, phi: torch.Tensor):
        # on a stream of its own (preparation waits for its own work only); the
        # handle is complete on return (its caustic guard is read) and serves
        # every plan of this N
        with torch.cuda.stream(self.prep_stream):
            return self._lib.PhiHandle(self._plan(0, N), phi.to(self.device), self.nsteps,
                                       with_min_det=True)

    def release_phi(self, handle):
        handle.close()

    def alloc_output(self, req: Request) -> torch.Tensor:
        return torch.empty_like(req.x)

    # batches ------------------------------------------------------------------------
    def start_clock(self) -> float:
        """Anchor the event clock: returns the host time of the anchor event."""
        ev = torch.cuda.Event(enable_timing=True)
        ev.record(torch.cuda.current_stream(self.device))
        ev.synchronize()
        self._anchor_t = time.perf_counter()
        self._anchor_ev = ev
        return self._anchor_t

    def launch(self, key, reqs: List[Request], handles, si: int):
        if self._anchor_ev is None:
            self.start_clock()
        N, C, _, _ = key
        for r in reqs:
            if r.x.dtype != self.dtype or not r.x.is_contiguous() or r.x.device != self.device:
                raise ValueError("request maps must be contiguous, on the serving device, in the "
                                 "serving dtype")
        cls = _batch_class(len(reqs), self.mb[N])
        ex = self._exec(si, key, cls)
        stream = self.streams[si]
        # inputs written on the caller's stream become visible to this one
        stream.wait_stream(torch.cuda.current_stream(self.device))
        ex.run(handles, [r.x.data_ptr() for r in reqs], [r.out.data_ptr() for r in reqs],
               stream.cuda_stream)
        ev = torch.cuda.Event(enable_timing=True)
        ev.record(stream)
        return ev

    # Wiener-filter lane ---------------------------------------------------------------
    def wf_bytes(self, req: WFRequest) -> int:
        """CG vectors (x, r, z, p, b: complex double), the operator's copy of p
        and the two plans' scratch (FP32 operator + FP64 residual), per
        right-hand side."""
        m = getattr(req.data, req.data._matrix_names()[0])
        nrhs = m.shape[0] if m.dim() == 3 else 1
        N, C = req.N, req.C
        P, H = N * N, N * (N // 2 + 1)
        return nrhs * C * (6 * 16 * H + 10 * 8 * P)

    def launch_wf(self, req: WFRequest, phi: torch.Tensor):
        """Start the solve (wiener_filter_process) on a host thread of its own
        that drives the lane's stream: the scheduler thread never spends time
        enqueueing a solve's ~170 launches per CG iteration, so LenseFlow
        batches keep being formed and launched while it runs (the library's
        calls and torch's launches release the interpreter lock)."""
        from .fields import FlatS0
        from .wiener_filter import wiener_filter_process
        if not self._warm:
            self.warmup()
        if self._anchor_ev is None:
            self.start_clock()
        if self._wf_stream is None:
            self._wf_stream = torch.cuda.Stream(self.device)
        self._wf_stream.wait_stream(torch.cuda.current_stream(self.device))
        Cf, Cn, M, B = req.ops
        phif = FlatS0(scalar_matrix=phi, theta_pix=self.theta_pix)
        tok = _WFToken(wiener_filter_process(req.data * 0, phif, req.data, Cf, Cn, M, B, n=req.n,
                                             tol=0.0, rtol=req.rtol, recompute_every=20,
                                             maxiter=req.maxiter, return_info=True))
        tok.thread = threading.Thread(target=self._run_wf, args=(tok,), daemon=True)
        self._wf_threads = [t for t in getattr(self, "_wf_threads", []) if t.is_alive()]
        self._wf_threads.append(tok.thread)
        # the lane's thread and the scheduler share the interpreter lock; a
        # short switch interval (0.2 ms instead of CPython's 5 ms) bounds how
        # long either waits for it while both run Python between launches
        if self._switch_interval is None:
            self._switch_interval = sys.getswitchinterval()
            sys.setswitchinterval(min(self._switch_interval, 2e-4))
        tok.thread.start()
        return tok

    def _run_wf(self, tok: "_WFToken"):
        """The lane's thread: run the process to completion on the lane's
        stream, then record the end event (read by finish_time)."""
        from .wiener_filter import run_process
        with torch.cuda.device(self.device), torch.cuda.stream(self._wf_stream):
            try:
                tok.result = run_process(tok.gen)
            except Exception as e:  # reported on the request; the server keeps running
                tok.error = f"{type(e).__name__}: {e}"
            end = torch.cuda.Event(enable_timing=True)
            end.record(self._wf_stream)
        tok.end = end
        tok.finished = True

    def wf_result(self, token, req: WFRequest) -> Optional[str]:
        if token.error is None:
            req.out, req.info = token.result
        return token.error

    def ready(self, token) -> bool:
        if isinstance(token, _WFToken):
            return token.finished and token.end.query()
        return token.query()

    def finish_time(self, token) -> float:
        if isinstance(token, _WFToken):
            token = token.end
        return self._anchor_t + self._anchor_ev.elapsed_time(token) * 1e-3

    def synchronize(self):
        torch.cuda.synchronize(self.device)

    def close(self):
        for t in getattr(self, "_wf_threads", []):
            t.join()
        if self._switch_interval is not None:
            sys.setswitchinterval(self._switch_interval)
            self._switch_interval = None
        torch.cuda.synchronize(self.device)
        for ex in self.execs.values():
            ex.close()
        self.execs.clear()
        for p in self.plans.values():
            p.close()
        self.plans.clear()


def poisson_schedule(n: int, rate: float, seed: int) -> np.ndarray:
    """Arrival offsets (s) of a Poisson stream: cumulative Exp(1/rate) gaps."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def mixed_workload(n: int, seed: int, sizes=(64, 128, 256, 512), weights=(0.2, 0.4, 0.3, 0.1),
                   nphi: int = 32):
    """Request descriptors of the mixed stream: (N, C, direction, phi_id) with
    N drawn with `weights`, T/QU and lense/unlense 50/50, phi_id ~ U{0..nphi-1}."""
    rng = np.random.default_rng(seed)
    Ns = rng.choice(np.asarray(sizes), size=n, p=np.asarray(weights) / np.sum(weights))
    Cs = rng.integers(1, 3, size=n)
    dirs = np.where(rng.random(n) < 0.5, FORWARD_LENSE, INVERSE_LENSE)
    phis = rng.integers(0, nphi, size=n)
    return [(int(a), int(b), int(c), int(d)) for a, b, c, d in zip(Ns, Cs, dirs, phis)]
