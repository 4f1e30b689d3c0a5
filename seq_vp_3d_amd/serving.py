"""Serving and batching forms: many registration pairs of uneven sizes sharing one MI355X.

A request carries a fixed and a moving image (host arrays) and an outer-step budget. The server admits requests
into slots -- each slot a worker thread with its own HIP stream -- while the sum of the admitted requests' HBM
footprints stays within the budget (first fit over the waiting queue in arrival order), so requests that arrive
while others are on the GPU start as soon as a slot and memory free up (continuous batching). An admitted
request keeps everything on the device until its certified result is returned: images, spline coefficients,
fields, the Barzilai-Borwein history and the metric. Its result is re-certified on the CPU by the interval
certificate (certify_map) before it is returned.

Every kernel is deterministic and a request's work is ordered on its own stream, so a result is bitwise the
same alone and in the mix (tests/test_gpu_serving.py checks it). The reference serves pairs one at a time from
one process (scripts/run_stage_3d_e48_pg.py:874-918).
"""
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np
import torch
import xxhash

from .certificate import CertificateConfig, certify_map
from .hotpath import PairState
from .problem import Q1Problem, make_grid_3d, native_weights
from .solver import PGConfig, SequentialConfig, solve_sequential


def footprint_bytes(shape_zyx):
    """HBM bytes one resident request needs (upper estimate): 17 node fields of 3 doubles (state, trial, reference,
    image gradients of the state and the trial, node terms, snapshot, direction field), 8 node scalars (images, warped, spline coefficients,
    certificate margins), 12 free vectors (theta, trial, gradient, direction, scale, diagonal, BB history,
    snapshot), the cell-local gradient buffer (48 doubles per cell) and 64 MiB of small buffers, plus 10%."""
    nz, ny, nx = shape_zyx
    n = nz * ny * nx
    c = (nz - 1) * (ny - 1) * (nx - 1)
    f = 3 * (nz - 2) * (ny - 2) * (nx - 2)
    total = 8 * (17 * 3 * n + 8 * n + 12 * f + 48 * c) + (64 << 20)
    return int(total * 1.1)


@dataclass
class Request:
    """One registration: images (host float64 arrays or CPU tensors -- pinned ones upload asynchronously -- of the
    same (nz, ny, nx) shape on a unit-spacing grid), the number of certified outer steps to run and whether the
    solver caps its steps by the device cubic JD root (T66)."""
    fixed: object
    moving: object
    max_outer: int = 10
    request_id: int = 0
    arrival: float = 0.
    recertify: bool = True
    jd_step_global: bool = True

    @property
    def shape(self):
        return tuple(self.fixed.shape)


@dataclass
class Result:
    request_id: int
    shape: tuple
    status: str
    outer_accepted: int
    energy: float
    theta_digest: str                 # XXH3-128 of the output free vector's bytes (bitwise identity of results)
    recertified: str | None
    min_jd_lower_bound: float | None
    arrival: float
    started: float
    finished: float
    seconds: dict = field(default_factory=dict)
    trajectory: list = field(default_factory=list)     # (reason, tau, trial energy) per line-search attempt

    @property
    def latency(self):
        return self.finished - self.arrival


def _upload(a, device):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.float64, non_blocking=True).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(device, non_blocking=True)


def _download(t):
    host = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    host.copy_(t, non_blocking=True)
    return host


def run_request(req, device, stream=None, recert_workers=8, batch=None, graphs=False):
    """Runs one request on `stream` (the current stream when None): upload, device spline, certified solver from
    u = 0, CPU re-certification of the output. With `batch` (a batch.BatchGroup) the trial-energy sweeps go through
    the group's multi-pair launches. `graphs`: the pair's trial and direction segments as captured HIP graphs (only
    where no other host thread uses the device, see Server). Times are time.perf_counter values."""
    started = time.perf_counter()
    ctx = torch.cuda.stream(stream) if stream is not None else torch.cuda.stream(torch.cuda.current_stream(device))
    secs = {}
    with ctx:
        nz, ny, nx = req.shape
        grid = make_grid_3d(nx=nx, ny=ny, nz=nz, hx=1., hy=1., hz=1.)
        fixed, moving = (_upload(a, device) for a in (req.fixed, req.moving))
        st = PairState(Q1Problem(grid, native_weights(grid.shape), 0., .05), fixed, moving, device, batch=batch,
                       graphs=graphs)
        st.set_state(theta=torch.zeros(st.nfree, dtype=torch.float64, device=device))
        t1 = time.perf_counter()
        secs["setup"] = t1 - started
        trajectory = []
        if batch is not None:
            batch.join()
        try:
            res = solve_sequential(st, SequentialConfig(), PGConfig(jd_step_global=req.jd_step_global),
                                   max_outer=req.max_outer,
                                   inner_observer=lambda r: trajectory.append((r["reason"], r["tau"], r["energy_trial"])))
        finally:
            if batch is not None:
                batch.leave()
        t2 = time.perf_counter()
        secs["solve"] = t2 - t1
        # downloads into page-locked buffers (torch's host cache reuses them across requests), one wait
        pinned = [_download(st.theta)] + ([_download(st.u)] if req.recertify else [])
        torch.cuda.current_stream(device).synchronize()
        theta = pinned[0].numpy()
        u = pinned[1].numpy() if req.recertify else None
        energy = res.energy.total
    digest = xxhash.xxh3_128_hexdigest(memoryview(theta).cast("B"))
    rec = None
    low = None
    if u is not None:
        cert = certify_map(u, grid, CertificateConfig(delta=0.), workers=recert_workers)
        rec, low = cert.status, cert.min_jd_lower_bound
    secs["recertify"] = time.perf_counter() - t2
    return Result(req.request_id, req.shape, res.status, res.outer_accepted, energy, digest, rec, low, req.arrival,
                  started, time.perf_counter(), secs, trajectory)


class Server:
    """Continuous-batching request loop on one device: `slots` worker threads, each with its own stream, admitting
    requests first-fit in arrival order while their footprints fit `hbm_budget` bytes.

    The slots issue their segments call by call (PairState(graphs=False)): while one host thread captures a HIP
    graph, the HIP runtime fails blocking copies and synchronisations issued by other threads with
    hipErrorStreamCaptureUnsupported ("operation not permitted when stream is capturing"), even in thread-local
    capture mode, and the slots synchronise all the time. Results are the same bits either way."""

    def __init__(self, device="cuda", slots=4, hbm_budget=None, recert_workers=4, batch=None):
        self.device = torch.device(device)
        self.batch = batch              # batch.BatchGroup shared by the slots (None: each pair launches alone)
        self.slots = int(slots)
        if hbm_budget is None:
            free, _total = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(free * .85)
        self.hbm_budget = int(hbm_budget)
        self.recert_workers = recert_workers
        self._cv = threading.Condition()
        self._queue = []                 # [(order, request, future, footprint)]
        self._order = 0
        self._used = 0
        self._closed = False
        self.admissions = []             # (time, request_id, footprint, bytes in use after admission, running)
        self._running = 0
        self._streams = [torch.cuda.Stream(device=self.device) for _ in range(self.slots)]
        self._threads = [threading.Thread(target=self._worker, args=(i,), name=f"seqvp-slot-{i}", daemon=True)
                         for i in range(self.slots)]
        for t in self._threads:
            t.start()

    def submit(self, req):
        """Queue a request; returns a Future of its Result (a MemoryError if it can never fit the budget)."""
        fut = Future()
        need = footprint_bytes(req.shape)
        if need > self.hbm_budget:
            fut.set_exception(MemoryError(f"request {req.request_id}: {need} B exceeds the budget {self.hbm_budget} B"))
            return fut
        with self._cv:
            self._queue.append((self._order, req, fut, need))
            self._order += 1
            self._cv.notify_all()
        return fut

    def _take(self):
        for i, (_, req, fut, need) in enumerate(self._queue):
            if self._used + need <= self.hbm_budget:
                del self._queue[i]
                return req, fut, need
        return None

    def _worker(self, slot):
        torch.cuda.set_device(self.device)
        stream = self._streams[slot]
        while True:
            with self._cv:
                while True:
                    got = self._take() if self._queue else None
                    if got is not None:
                        break
                    if self._closed:
                        return
                    self._cv.wait(timeout=.05)
                req, fut, need = got
                self._used += need
                self._running += 1
                self.admissions.append((time.perf_counter(), req.request_id, need, self._used, self._running))
            try:
                fut.set_result(run_request(req, self.device, stream, self.recert_workers, self.batch, graphs=False))
            except BaseException as exc:          # a failed request never takes the server down
                fut.set_exception(exc)
            finally:
                with self._cv:
                    self._used -= need
                    self._running -= 1
                    self._cv.notify_all()

    def close(self):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()


def serve(requests, offsets=None, device="cuda", slots=4, hbm_budget=None, recert_workers=4, batch=None):
    """Submit the requests at their arrival offsets (seconds from the start; None = all at once, the batching
    form) and collect their results. `batch`: a batch.BatchGroup for multi-pair trial launches. Returns (results
    in request order, wall seconds, server)."""
    server = Server(device, slots=slots, hbm_budget=hbm_budget, recert_workers=recert_workers, batch=batch)
    t0 = time.perf_counter()
    futures = []
    for i, req in enumerate(requests):
        if offsets is not None:
            wait = offsets[i] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        req.arrival = time.perf_counter()
        futures.append(server.submit(req))
    results = [f.result() for f in futures]
    wall = time.perf_counter() - t0
    server.close()
    return results, wall, server


def serial(requests, device="cuda", recert_workers=4):
    """The plain path: one request at a time on the current stream. Returns (results, wall seconds)."""
    t0 = time.perf_counter()
    out = []
    for req in requests:
        req.arrival = time.perf_counter()
        out.append(run_request(req, torch.device(device), None, recert_workers))
    return out, time.perf_counter() - t0


def poisson_offsets(n, rate, seed):
    """Arrival offsets (s) of n requests of a Poisson process of the given rate (exponential gaps)."""
    return np.cumsum(np.random.default_rng(seed).exponential(1. / rate, size=n))


def latency_summary(results):
    lat = np.array([r.latency for r in results])
    return dict(p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)), mean=float(lat.mean()),
                max=float(lat.max()))
