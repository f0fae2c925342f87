"""Bounded dispatcher over the native slots (mirror of
emitter/src/main/java/br/unb/cic/tdp/base/GPUSortingSearch.java:40-264).

A bounded pending queue of capacity ``total_slots + queue_capacity``
(:46-52) feeds one worker thread per slot (:123-162); each worker makes the
blocking native call for its slot.  ``search`` converts the returned
positions to symbols by replaying the block swaps on a copy of beta
(:229-239).  With ``served=True`` the workers hand their requests to the
served loop instead (continuous batching in one persistent kernel); the
Python API is the same.
"""
from __future__ import annotations

import os
import queue
import threading
from concurrent.futures import Future
from typing import List, Optional, Sequence, Tuple

from . import native


def _prop(name: str, default: str = "") -> str:
    # the reference reads -Dtdp.* system properties; here TDP_* environment variables
    return os.environ.get("TDP_" + name, default).strip()


class GPUSortingSearch:
    _instance: Optional["GPUSortingSearch"] = None
    _lock = threading.Lock()

    def __init__(self, devices: Sequence[int], slots_per_device: int, queue_bytes_budget: int,
                 queue_capacity: Optional[int] = None, served: bool = False,
                 served_slots: int = 0):
        if queue_bytes_budget <= 0:
            raise ValueError("queueBytesBudget must be positive")
        self.devices = [int(d) for d in devices]
        if not self.devices or min(self.devices) < 0 or \
                len(set(self.devices)) != len(self.devices):
            raise ValueError("the device list needs distinct non-negative ordinals, got %r"
                             % (list(devices),))
        if slots_per_device < 1:
            raise ValueError("slots per device must be at least 1, got %d" % slots_per_device)
        self.total_slots = len(self.devices) * slots_per_device
        if queue_capacity is None:
            v = _prop("GPU_QUEUE_CAPACITY")
            queue_capacity = int(v) if v else 0
        if queue_capacity < 0:
            raise ValueError("tdp.gpuQueueCapacity must not be negative")
        self.queue_capacity = queue_capacity
        self.pending_capacity = self.total_slots + queue_capacity
        self._pending: "queue.Queue" = queue.Queue(maxsize=self.pending_capacity)
        # finished searches, successful and failed (the Java side keeps its
        # own windowed statistics around the native call)
        self._count_lock = threading.Lock()
        self.completed = 0
        self.failed = 0
        self.served = served
        if served:
            if len(self.devices) != 1:
                raise ValueError("the served loop runs on one device")
            native.serve_start(device=self.devices[0],
                               slots=served_slots or max(self.total_slots, 64),
                               queue_capacity=max(queue_capacity, 1),
                               slot_bytes=queue_bytes_budget)
        else:
            # slot s runs on devices[s % len(devices)] (td16_init_slots)
            native.init_slots(self.total_slots, self.devices, queue_bytes_budget)
        self._threads = []
        for slot in range(self.total_slots):
            t = threading.Thread(target=self._run_slot, args=(slot,), daemon=True,
                                 name="td16-slot-%d" % slot)
            t.start()
            self._threads.append(t)

    # -- singleton (GPUSortingSearch.java:96-113) -------------------------
    @classmethod
    def init(cls, devices: Sequence[int], slots_per_device: int, queue_bytes_budget: int,
             **kw) -> None:
        with cls._lock:
            if cls._instance is None:
                cls._instance = GPUSortingSearch(devices, slots_per_device, queue_bytes_budget,
                                                 **kw)

    @classmethod
    def get_instance(cls) -> "GPUSortingSearch":
        if cls._instance is None:
            raise RuntimeError("GPUSortingSearch has not been initialized. Call init() first.")
        return cls._instance

    @classmethod
    def reset(cls) -> None:
        """Stops the workers and releases the native slots (tests only)."""
        with cls._lock:
            inst = cls._instance
            cls._instance = None
        if inst is not None:
            for _ in inst._threads:
                inst._pending.put(None)
            for t in inst._threads:
                t.join(timeout=30)
            if inst.served:
                native.serve_stop()
            native.shutdown()

    # -- workers (GPUSortingSearch.java:134-162) ---------------------------
    def _run_slot(self, slot: int) -> None:
        while True:
            job = self._pending.get()
            if job is None:
                return
            beta, omega, initial_even, min_rate, max_moves, full_sorting, fut = job
            failed = True
            try:
                if self.served:
                    req = native.make_request(beta, omega, initial_even, min_rate, max_moves,
                                              full_sorting)
                    res = native.serve_search(req)
                    result = res.witness()
                else:
                    result = native.search_positions(slot, beta, omega, initial_even, min_rate,
                                                     max_moves, full_sorting)
                failed = False
                fut.set_result(result)
            except BaseException as e:  # the job completes exceptionally (:150-152)
                fut.set_exception(e)
            finally:
                with self._count_lock:
                    self.completed += 0 if failed else 1
                    self.failed += 1 if failed else 0

    # -- search (GPUSortingSearch.java:203-240) ----------------------------
    def search(self, beta: Sequence[int], omega: Sequence[int], initial_even: int,
               min_rate: float, max_moves: int, full_sorting: bool = True
               ) -> List[Tuple[int, int, int]]:
        fut: Future = Future()
        self._pending.put((list(beta), list(omega), initial_even, min_rate, max_moves,
                           full_sorting, fut))
        positions = fut.result()
        if not positions:
            return []
        return native.positions_to_symbols(beta, positions)
