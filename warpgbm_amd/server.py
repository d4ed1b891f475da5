"""Continuous-batching predict server over a resident heap forest.

Requests (int8 bin-code matrices of any row count) are queued; a worker
thread cuts them into pieces of at most ``piece_rows`` rows and packs pieces
into traversal launches ("batches") of at most ``max_batch_rows`` rows,
shortest remaining request first, while the device memory they need fits a
budget.  Up to ``slots`` batches are in flight on separate HIP streams, so
the staging and host-to-device copy of one batch overlap the traversal of
the previous one.  Requests that arrive while batches are running join the
next batch as soon as a slot and budget free up; a large request no longer
holds up small ones behind it (its pieces yield to them), and its own pieces
pipeline through the slots.

Each row's score depends only on its own codes and the forest (the serving
kernel sums trees in a fixed order), so a request's result is bit-identical
whether it is served alone, in pieces, or mixed with others.
"""
from __future__ import annotations

import collections
import sys
import threading
import time
from concurrent.futures import Future

import numpy as np
import torch


class _Request:
    __slots__ = ("codes", "future", "t_submit", "t_launch", "rows", "next_row", "done_rows", "out", "seq")

    def __init__(self, codes, future, n_classes, seq):
        self.codes = codes
        self.future = future
        self.t_submit = time.perf_counter()
        self.t_launch = 0.0
        self.rows = codes.shape[0]
        self.next_row = 0   # rows handed to batches so far
        self.done_rows = 0  # rows whose scores are back
        self.out = np.empty((self.rows, n_classes), dtype=np.float32)
        self.seq = seq


class _Slot:
    def __init__(self, forest, max_rows, device):
        F, K = forest.n_feat, forest.n_classes
        self.stream = torch.cuda.Stream(device=device)
        self.h_codes = torch.empty((max_rows, F), dtype=torch.int8, pin_memory=True)
        self.d_codes = torch.empty((max_rows, F), dtype=torch.int8, device=device)
        self.d_out = torch.empty((max_rows, K), dtype=torch.float32, device=device)
        self.h_out = torch.empty((max_rows, K), dtype=torch.float32, pin_memory=True)
        # Touch every page now: first writes to fresh pinned memory run at ~3 GB/s
        # (page mapping) against ~30 GB/s afterwards, and must not hit requests.
        self.h_codes.numpy().fill(0)
        self.h_out.numpy().fill(0)
        need = forest.workspace(max_rows).numel()
        self.ws = torch.empty(need, dtype=torch.uint8, device=device)
        self.event = None
        self.reqs = []
        self.rows = 0
        self.bytes = 0


class PredictServer:
    """Serve predictions of a compiled ``HeapForest``.

    memory_budget: device bytes the in-flight pieces may hold (codes,
    outputs and reduction workspace per row); a request needing more than
    the whole budget is rejected.  max_batch_rows caps one launch and
    piece_rows one piece of a request (default: a whole launch, so a large
    request goes through in the fewest launches; smaller requests still
    take the next free slot ahead of its remaining pieces).

    Latency tails: the worker is a Python thread, so a full cyclic garbage
    collection anywhere in the process (which walks every live object, e.g.
    a fitted model's nested tree dicts) stalls it.  Serving processes should
    call ``gc.freeze()`` once their long-lived objects exist.  The worker
    also competes with client threads for the interpreter lock, which a
    thread keeps for up to ``sys.getswitchinterval()`` (5 ms by default):
    ``switch_interval`` (seconds) lowers that process-wide setting while the
    server runs (None leaves it alone).  Measured on a 1000-tree depth-8
    forest with log-uniform 1..100k-row requests at ~4300 requests/s: p99
    3.9 ms at 5 ms, 2.5 ms at 0.2 ms.
    """

    def __init__(self, forest, memory_budget=1 << 30, max_batch_rows=32768, slots=3, piece_rows=None,
                 switch_interval=2e-4):
        self.forest = forest
        dev = torch.device(forest.device)
        if dev.index is None:
            dev = torch.device(dev.type, torch.cuda.current_device())
        self.device = dev
        self.max_batch_rows = int(max_batch_rows)
        self.piece_rows = (self.max_batch_rows if piece_rows is None
                           else max(1, min(int(piece_rows), self.max_batch_rows)))
        self.budget = int(memory_budget)
        K = forest.n_classes
        self.bytes_per_row = forest.n_feat + 8 * K + 4 * K * forest.n_blocks
        self._slots = [_Slot(forest, self.max_batch_rows, self.device) for _ in range(slots)]
        self._warm_slots()
        self._free = collections.deque(self._slots)
        self._busy = collections.deque()
        self._queue = []  # requests with rows not yet handed to a batch
        self._cv = threading.Condition()
        self._in_flight_bytes = 0
        self._seq = 0
        self._stop = False
        self._thread = threading.Thread(target=self._loop, name="wgbm-predict-server", daemon=True)
        self._saved_switch = None
        if switch_interval is not None and sys.getswitchinterval() > switch_interval:
            self._saved_switch = sys.getswitchinterval()
            sys.setswitchinterval(switch_interval)
        self.latencies_ms = []
        self.queue_ms = []  # submit -> first launch of each completed request
        self.batch_log = None  # set to a list to record (slot, rows, t_launch, t_copied, t_done) per batch
        self.batches = 0
        self._ready = threading.Event()
        self._start_error = None
        self._thread.start()
        self._ready.wait()
        if self._start_error is not None:
            self.close()
            raise RuntimeError("the predict server's worker failed to start") from self._start_error

    def _warm_slots(self):
        """A launch and large copies on every slot's stream: the runtime
        creates a stream's kernel queue and its copy-engine queue at their
        first use (tens of ms each), which must not land on the first
        requests.  Small and large copies take different engines, so the
        warm-up copies a few MB each way."""
        F = self.forest.n_feat
        for s in self._slots:
            n = min(s.h_codes.shape[0], max(1, (4 << 20) // max(F, 1)))
            s.h_codes[:n].zero_()
            with torch.cuda.stream(s.stream):
                for m in (1, n):
                    s.d_codes[:m].copy_(s.h_codes[:m], non_blocking=True)
                    self.forest.predict_into(s.d_codes[:m], s.d_out[:m], stream=s.stream.cuda_stream,
                                             workspace=s.ws)
                    s.h_out[:m].copy_(s.d_out[:m], non_blocking=True)
            s.stream.synchronize()

    # ---------------------------------------------------------------- client
    def submit(self, codes: np.ndarray) -> Future:
        """Queue one request (int8 [rows, n_feat]); the future yields float32
        [rows, n_classes] scores."""
        codes = np.ascontiguousarray(codes, dtype=np.int8)
        if codes.ndim != 2 or codes.shape[1] != self.forest.n_feat:
            raise ValueError(f"request must be int8 [rows, {self.forest.n_feat}]")
        if codes.shape[0] * self.bytes_per_row > self.budget:
            raise ValueError("request exceeds the server's memory budget")
        fut = Future()
        if codes.shape[0] == 0:
            fut.set_result(np.empty((0, self.forest.n_classes), np.float32))
            return fut
        with self._cv:
            # after close() (or a worker that died) nothing would ever serve it
            if self._stop or not self._thread.is_alive():
                raise RuntimeError("the predict server is closed")
            self._queue.append(_Request(codes, fut, self.forest.n_classes, self._seq))
            self._seq += 1
            self._cv.notify()
        return fut

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify()
        self._thread.join()
        if getattr(self, "_saved_switch", None) is not None:
            sys.setswitchinterval(self._saved_switch)
            self._saved_switch = None

    # ---------------------------------------------------------------- worker
    def _warm_worker(self):
        """Give this thread torch's intra-op thread count and stage a full
        launch into a slot's pinned buffer a few times.  A new thread's
        large host copies otherwise run on a smaller team (measured on a
        13 MB launch: 0.45 ms against 0.12 ms), and the team is created on
        the first parallel copy (~10 ms), which must not land on the first
        request."""
        torch.set_num_threads(torch.get_num_threads())
        s = self._slots[0]
        src = torch.from_numpy(np.zeros(tuple(s.h_codes.shape), dtype=np.int8))
        for _ in range(4):
            s.h_codes.copy_(src)

    def _loop(self):
        try:
            torch.cuda.set_device(self.device)
            self._warm_worker()
            self._ready.set()
            self._serve()
        except BaseException as e:  # fail every pending request instead of hanging its caller
            if not self._ready.is_set():
                self._start_error = e
                self._ready.set()
                return
            with self._cv:
                pending = [r for s in self._busy for r, _, _ in s.reqs] + list(self._queue)
                self._queue.clear()
            for r in pending:
                if not r.future.done():
                    r.future.set_exception(e)
            raise

    def _serve(self):
        while True:
            with self._cv:
                if self._stop and not self._busy and not self._queue:
                    return
                if not self._busy and not (self._queue and self._free):
                    self._cv.wait(timeout=0.05)
                    continue
                batch = self._admit_locked()
            if batch is not None:
                self._launch(*batch)
            done = self._complete_ready()
            if batch is None and not done:
                time.sleep(20e-6)

    def _admit_locked(self):
        """Pieces for one batch: the request with the fewest rows left goes
        first (arrival order among equals), each contributing at most
        piece_rows rows, while the batch's rows and the memory budget allow."""
        if not self._queue or not self._free:
            return None
        pieces, rows, nbytes = [], 0, 0
        order = sorted(self._queue, key=lambda r: (r.rows - r.next_row, r.seq))
        for r in order:
            while r.next_row < r.rows and rows < self.max_batch_rows:
                k = min(r.rows - r.next_row, self.piece_rows, self.max_batch_rows - rows)
                b = k * self.bytes_per_row
                if self._in_flight_bytes + nbytes + b > self.budget:
                    break
                pieces.append((r, r.next_row, r.next_row + k))
                r.next_row += k
                rows += k
                nbytes += b
            if rows >= self.max_batch_rows or (r.next_row < r.rows):
                break  # batch full, or the budget stopped this request's next piece
        if not pieces:
            return None
        self._queue = [r for r in self._queue if r.next_row < r.rows]
        self._in_flight_bytes += nbytes
        return self._free.popleft(), pieces, rows, nbytes

    def _launch(self, slot, pieces, rows, nbytes):
        t = time.perf_counter()
        o = 0
        hc = slot.h_codes
        for r, a, b in pieces:
            if r.t_launch == 0.0:
                r.t_launch = t
            # staging copy into pinned memory (torch copies large blocks on several threads)
            hc[o:o + b - a].copy_(torch.from_numpy(r.codes[a:b]))
            o += b - a
        with torch.cuda.stream(slot.stream):
            slot.d_codes[:rows].copy_(slot.h_codes[:rows], non_blocking=True)
            self.forest.predict_into(slot.d_codes[:rows], slot.d_out[:rows], stream=slot.stream.cuda_stream,
                                     workspace=slot.ws)
            slot.h_out[:rows].copy_(slot.d_out[:rows], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(slot.stream)
        slot.event, slot.reqs, slot.rows, slot.bytes = ev, pieces, rows, nbytes
        slot.t_launch, slot.t_copied = t, time.perf_counter()
        self.batches += 1
        self._busy.append(slot)

    def _complete_ready(self):
        """Deliver every finished batch (in launch order); True if any."""
        done = False
        while self._busy and self._busy[0].event.query():
            slot = self._busy.popleft()
            if self.batch_log is not None:
                self.batch_log.append((self._slots.index(slot), slot.rows, slot.t_launch, slot.t_copied,
                                       time.perf_counter()))
            out = slot.h_out.numpy()
            o = 0
            finished = []
            for r, a, b in slot.reqs:
                r.out[a:b] = out[o:o + b - a]
                o += b - a
                r.done_rows += b - a
                if r.done_rows == r.rows:
                    finished.append(r)
            with self._cv:
                self._in_flight_bytes -= slot.bytes
                slot.reqs = []
                self._free.append(slot)
            now = time.perf_counter()
            for r in finished:
                self.latencies_ms.append((now - r.t_submit) * 1e3)
                self.queue_ms.append((r.t_launch - r.t_submit) * 1e3)
                r.future.set_result(r.out)
            done = True
        return done
