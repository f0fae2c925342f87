"""Slot -> device mapping (mirror of
emitter/src/main/java/br/unb/cic/tdp/base/GpuDeviceLayout.java:15-96)."""
from __future__ import annotations

from typing import Sequence

INT_MAX = 2**31 - 1


class GpuDeviceLayout:
    """Validated, equal-capacity mapping from logical search slots to devices."""

    def __init__(self, devices: Sequence[int], slots_per_device: int):
        if devices is None or len(devices) == 0:
            raise ValueError("devices must contain at least one device ordinal")
        if slots_per_device <= 0:
            raise ValueError("slotsPerDevice must be positive")
        seen = set()
        for d in devices:
            if d < 0:
                raise ValueError("devices must contain only non-negative device ordinals")
            if d in seen:
                raise ValueError(f"devices must contain distinct device ordinals; duplicate: {d}")
            seen.add(d)
        total = len(devices) * slots_per_device
        if total > INT_MAX:  # Math.multiplyExact overflow in the reference
            raise ValueError("total GPU slot count is too large")
        self._devices = list(devices)
        self._slots_per_device = slots_per_device
        self._total = total

    @staticmethod
    def parse_devices(value: str):
        if value is None or value.strip() == "":
            raise ValueError("GPU device list must not be blank")
        out = []
        for token in value.split(","):
            token = token.strip()
            if token == "":
                raise ValueError("GPU device list contains an empty ordinal")
            try:
                out.append(int(token, 10))
            except ValueError as e:
                raise ValueError(f"GPU device list contains a non-integer ordinal: {value}") from e
        return out

    def devices(self):
        return list(self._devices)

    def device_count(self) -> int:
        return len(self._devices)

    def slots_per_device(self) -> int:
        return self._slots_per_device

    def total_slots(self) -> int:
        return self._total

    def device_for_slot(self, slot: int) -> int:
        self._check(slot)
        return self._devices[slot % len(self._devices)]

    def device_index_for_slot(self, slot: int) -> int:
        self._check(slot)
        return slot % len(self._devices)

    def _check(self, slot: int) -> None:
        if slot < 0 or slot >= self._total:
            raise ValueError(f"GPU slot is outside the configured range: {slot}")
