"""Batching form: the trial-energy sweeps of many pairs of uneven sizes in one persistent launch.

Each pair of a batch runs its own certified solver on its own host thread and HIP stream (serving.Server); only the
dominant sweep of a line-search attempt -- the trial's soft and closed27 barrier energies against the accepted
state (seqvp_q1_trial_energy) -- goes through a BatchGroup. The group gathers the pairs that reach that sweep
together (all running pairs, or whoever has arrived when the oldest has waited `max_wait` seconds), makes its
stream wait for each pair's stream, enqueues one seqvp_q1_trial_energy_batch launch over the work list of all their
cells, and makes each pair's stream wait for it. Every pair's energies are the bits of its own single-pair sweep,
whatever the batch, so a result does not depend on which pairs it shared launches with.

The launch shape -- `tile` 256-cell blocks per work item and `grid` persistent workgroups -- is picked by
LaunchTuner, an online autotuner: per work-size class it times each candidate on live launches (all candidates give
the same bits) and keeps the fastest. The reference hand-tunes its launches per kernel (TUNED_LAUNCH,
jd_curl/three_d/backends/triton_q1.py:30-39); here the table is measured per batch size class on the GPU it runs on
and can be saved and reloaded (save / load).
"""
import ctypes
import json
import math
import threading
import time

import torch

from . import _lib

TRIAL_BATCH_MAX = 32           # pairs per launch (the kernel's argument table)
OP_TRIAL = 10


class TrialBatchEntry(ctypes.Structure):
    """SeqvpTrialBatchEntry (include/seqvp.h)."""
    _fields_ = [("u", ctypes.c_void_p), ("ref", ctypes.c_void_p), ("base", ctypes.c_void_p), ("k", ctypes.c_void_p),
                ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64), ("out", ctypes.c_void_p),
                ("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_size_t)]


def entry(q1, u, ref, base, out, workspace):
    """The batch entry of one pair's trial energy (FusedQ1 q1; node fields u, ref, base; 7-double out)."""
    q1._check(u, ref, base)
    nz, ny, nx = q1.shape
    need = _lib.workspace_bytes(OP_TRIAL, q1.shape)
    if workspace.numel() < need:
        raise ValueError("workspace smaller than the trial sweep's")
    return TrialBatchEntry(u.data_ptr(), ref.data_ptr(), base.data_ptr(), q1.k.data_ptr(), nx, ny, nz,
                           out.data_ptr(), workspace.data_ptr(), workspace.numel())


def blocks_of(shape):
    nz, ny, nx = shape
    return -(-((nx - 1) * (ny - 1) * (nz - 1)) // 256)


def trial_energy_batch(entries, tile, grid, stream):
    """One seqvp_q1_trial_energy_batch launch over `entries` (a list of TrialBatchEntry) on `stream`."""
    n = len(entries)
    if not 1 <= n <= TRIAL_BATCH_MAX:
        raise ValueError(f"1..{TRIAL_BATCH_MAX} pairs per launch")
    arr = (TrialBatchEntry * n)(*entries)
    _lib.call("seqvp_q1_trial_energy_batch", arr, n, int(tile), int(grid), stream)


# (tile, grid): blocks per work item, persistent workgroups (0 = one workgroup per item). MI355X holds 2 workgroups
# of the sweep per CU (256 VGPRs, 74 KB of LDS), 512 in all.
CANDIDATES = ((1, 0), (2, 0), (4, 0), (1, 512), (2, 512), (1, 1024), (4, 1024))


class LaunchTuner:
    """Online autotuner of the batched sweep's (tile, grid) per work-size class (log2 of the launch's blocks)."""

    def __init__(self, candidates=CANDIDATES, trials=3, table=None):
        self.candidates = tuple(tuple(c) for c in candidates)
        self.trials = int(trials)
        self.best = {int(k): tuple(v) for k, v in (table or {}).items()}
        self._samples = {}
        self._pending = []
        self._lock = threading.Lock()

    @staticmethod
    def size_class(blocks):
        return int(round(math.log2(max(int(blocks), 1))))

    def choose(self, blocks):
        """(class, (tile, grid), timed): the kept shape of the class, else the candidate with the fewest timings."""
        cls = self.size_class(blocks)
        with self._lock:
            self._harvest()
            if cls in self.best:
                return cls, self.best[cls], False
            samples = self._samples.setdefault(cls, {c: [] for c in self.candidates})
            waiting = {}
            for c_cls, cand, _, _ in self._pending:
                if c_cls == cls:
                    waiting[cand] = waiting.get(cand, 0) + 1
            cand = min(self.candidates, key=lambda c: len(samples[c]) + waiting.get(c, 0))
            return cls, cand, True

    def record(self, cls, cand, start, end):
        with self._lock:
            self._pending.append((cls, cand, start, end))

    def _harvest(self):
        keep = []
        for cls, cand, start, end in self._pending:
            if end.query():
                self._samples.setdefault(cls, {c: [] for c in self.candidates})[cand].append(start.elapsed_time(end))
            else:
                keep.append((cls, cand, start, end))
        self._pending = keep
        for cls, samples in self._samples.items():
            if cls not in self.best and all(len(v) >= self.trials for v in samples.values()):
                med = {c: sorted(v)[len(v) // 2] for c, v in samples.items()}
                self.best[cls] = min(med, key=med.get)

    def report(self):
        """{class: {"best": (tile, grid) or None, "median_ms": {candidate: ms}}} of the timings so far."""
        with self._lock:
            self._harvest()
            out = {}
            for cls, samples in sorted(self._samples.items()):
                med = {f"{c[0]}x{c[1]}": round(sorted(v)[len(v) // 2], 4) for c, v in samples.items() if v}
                out[cls] = dict(best=self.best.get(cls), median_ms=med)
            return out

    def save(self, path):
        with open(path, "w") as f:
            json.dump({str(k): list(v) for k, v in sorted(self.best.items())}, f, indent=2)

    @classmethod
    def load(cls, path, **kw):
        with open(path) as f:
            return cls(table=json.load(f), **kw)


class BatchGroup:
    """Rendezvous of the running pairs of a batch for their trial-energy sweeps (see the module docstring)."""

    def __init__(self, device="cuda", tuner=None, max_wait=.002):
        self.device = torch.device(device)
        self.stream = torch.cuda.Stream(device=self.device)
        self.tuner = tuner if tuner is not None else LaunchTuner()
        self.max_wait = float(max_wait)
        self._cv = threading.Condition()
        self._members = 0
        self._pending = []
        self.launches = 0
        self.pairs_launched = 0

    def join(self):
        with self._cv:
            self._members += 1

    def leave(self):
        with self._cv:
            self._members -= 1
            if self._pending and len(self._pending) >= self._members:
                self._launch()

    def trial_energy(self, st, out):
        """st's trial energies (u_t against ref, base u; PairState st) into `out` (7 device doubles) through the
        next batched launch. Returns once that launch is enqueued and st's current stream waits for it."""
        stream = torch.cuda.current_stream(st.device)
        ready = torch.cuda.Event()
        ready.record(stream)
        req = dict(entry=entry(st.q1, st.u_t, st.ref, st.u, out, st.q1.ws.buf), stream=stream, ready=ready,
                   blocks=blocks_of(st.shape), done=False)
        with self._cv:
            self._pending.append(req)
            if len(self._pending) >= self._members:
                self._launch()
            else:
                deadline = time.perf_counter() + self.max_wait
                while not req["done"]:
                    left = deadline - time.perf_counter()
                    if left <= 0:
                        self._launch()
                        break
                    self._cv.wait(left)

    def _launch(self):
        """Launches every pending request (under the lock), TRIAL_BATCH_MAX pairs per launch."""
        while self._pending:
            reqs, self._pending = self._pending[:TRIAL_BATCH_MAX], self._pending[TRIAL_BATCH_MAX:]
            for r in reqs:
                self.stream.wait_event(r["ready"])
            blocks = sum(r["blocks"] for r in reqs)
            cls, (tile, grid), timed = self.tuner.choose(blocks)
            handle = ctypes.c_void_p(self.stream.cuda_stream)
            if timed:
                start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                start.record(self.stream)
            trial_energy_batch([r["entry"] for r in reqs], tile, grid, handle)
            if timed:
                end.record(self.stream)
                self.tuner.record(cls, (tile, grid), start, end)
            done = torch.cuda.Event()
            done.record(self.stream)
            for r in reqs:
                r["stream"].wait_event(done)
                r["done"] = True
            self.launches += 1
            self.pairs_launched += len(reqs)
        self._cv.notify_all()
