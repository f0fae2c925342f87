"""Windowed slot statistics (mirror of
emitter/src/main/java/br/unb/cic/tdp/base/GpuSearchStats.java:63-131)."""
from __future__ import annotations

import threading
from dataclasses import dataclass, field
from typing import List

from .device_layout import GpuDeviceLayout


@dataclass(frozen=True)
class DeviceSnapshot:
    device: int
    started: int
    completed: int
    failed: int
    queue_wait_nanos: int
    busy_nanos: int


@dataclass(frozen=True)
class Snapshot:
    elapsed_nanos: int
    started: int
    completed: int
    failed: int
    queue_wait_nanos: int
    busy_nanos: int
    total_slots: int
    devices: List[DeviceSnapshot] = field(default_factory=list)


class GpuSearchStats:
    """All updates and snapshots are serialised so that counts and durations
    cannot be split across adjacent reporting windows."""

    def __init__(self, layout: GpuDeviceLayout, started_nanos: int):
        self._layout = layout
        self._lock = threading.Lock()
        self._window_start = started_nanos
        nd = layout.device_count()
        self._active_since = [-1] * layout.total_slots()
        self._started = [0] * nd
        self._completed = [0] * nd
        self._failed = [0] * nd
        self._wait = [0] * nd
        self._busy = [0] * nd

    def acquired(self, slot: int, queued_at: int, acquired_at: int) -> None:
        with self._lock:
            if self._active_since[slot] >= 0:
                raise RuntimeError(f"GPU slot {slot} was acquired while already active")
            d = self._layout.device_index_for_slot(slot)
            self._started[d] += 1
            self._wait[d] += max(0, acquired_at - queued_at)
            self._active_since[slot] = acquired_at

    def released(self, slot: int, released_at: int, failed: bool) -> None:
        with self._lock:
            since = self._active_since[slot]
            if since < 0:
                raise RuntimeError(f"GPU slot {slot} was released while inactive")
            d = self._layout.device_index_for_slot(slot)
            self._busy[d] += max(0, released_at - since)
            self._completed[d] += 1
            if failed:
                self._failed[d] += 1
            self._active_since[slot] = -1

    def snapshot(self, now: int) -> Snapshot:
        with self._lock:
            elapsed = max(1, now - self._window_start)
            for slot, since in enumerate(self._active_since):
                if since >= 0:
                    d = self._layout.device_index_for_slot(slot)
                    self._busy[d] += max(0, now - since)
                    self._active_since[slot] = now
            devs = self._layout.devices()
            snaps = [DeviceSnapshot(devs[i], self._started[i], self._completed[i], self._failed[i],
                                    self._wait[i], self._busy[i]) for i in range(len(devs))]
            snap = Snapshot(elapsed, sum(self._started), sum(self._completed), sum(self._failed),
                            sum(self._wait), sum(self._busy), self._layout.total_slots(), snaps)
            nd = len(devs)
            self._started = [0] * nd
            self._completed = [0] * nd
            self._failed = [0] * nd
            self._wait = [0] * nd
            self._busy = [0] * nd
            self._window_start = now
            return snap
