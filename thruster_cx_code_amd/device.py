"""Device plumbing: torch-ROCm buffers and raw pointers for the C ABI."""
import ctypes as C

import torch

from . import _lib


def require_gpu():
    """The backend has no CPU fallback: fail loudly without a HIP device."""
    if not torch.cuda.is_available():
        raise RuntimeError("thruster_cx_code_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    _lib.load()


def ptr(t):
    """Raw device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


# torch's direct bindings for the current device and its current raw stream
# (what torch's own generated code calls); torch.cuda.current_stream() costs
# several microseconds of Python per call, and a step makes a dozen calls
_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_cur_device = getattr(torch._C, "_cuda_getDevice", None)


def stream_ptr(stream=None):
    """hipStream_t of `stream`, or of the current torch stream."""
    if stream is None and _raw_stream is not None and _cur_device is not None:
        return C.c_void_p(_raw_stream(_cur_device()))
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class EventRing:
    """Reusable events for asynchronous readbacks, handed out in turn; a
    caller keeps fewer than `len(events) - 1` of them pending."""

    def __init__(self, n=64):
        self.events = [torch.cuda.Event() for _ in range(n)]
        self.i = 0

    def __len__(self):
        return len(self.events)

    def next(self):
        ev = self.events[self.i]
        self.i = (self.i + 1) % len(self.events)
        return ev


def empty(n, dtype, device):
    return torch.empty(int(n), dtype=dtype, device=device)


def zeros(n, dtype, device):
    return torch.zeros(int(n), dtype=dtype, device=device)


class SpeciesBuffers:
    """Structure-of-arrays particle storage of one species (capacity `cap`)
    with a device count.  Ions carry a uint32 tag: persistent id | isCEX<<31."""

    def __init__(self, cap, device, with_tag):
        self.cap = int(cap)
        self.device = device
        self.x = empty(cap, torch.float32, device)
        self.y = empty(cap, torch.float32, device)
        self.vx = empty(cap, torch.float32, device)
        self.vy = empty(cap, torch.float32, device)
        self.vz = empty(cap, torch.float32, device)
        self.tag = empty(cap, torch.int32, device) if with_tag else None
        self.count = zeros(1, torch.int32, device)
        self._struct = None
        self._ptrs = None

    def ptrs(self):
        """(x, y, vx, vy, vz, count) raw pointers (cached until `grow`)."""
        if self._ptrs is None:
            self._ptrs = (ptr(self.x), ptr(self.y), ptr(self.vx), ptr(self.vy), ptr(self.vz), ptr(self.count))
        return self._ptrs

    def struct(self):
        """The bemcs_species_t of these buffers (cached until `grow`; callers
        copy it or pass it by const pointer)."""
        if self._struct is None:
            self._struct = _lib.Species(ptr(self.x), ptr(self.y), ptr(self.vx), ptr(self.vy), ptr(self.vz),
                                        ptr(self.tag), ptr(self.count), self.cap)
        return self._struct

    def arrays(self):
        out = [self.x, self.y, self.vx, self.vy, self.vz]
        if self.tag is not None:
            out.append(self.tag)
        return out

    def grow(self, new_cap, n_valid):
        """Reallocate with a larger capacity, keeping the first n_valid entries."""
        new_cap = int(new_cap)
        for name in ("x", "y", "vx", "vy", "vz", "tag"):
            old = getattr(self, name)
            if old is None:
                continue
            new = torch.empty(new_cap, dtype=old.dtype, device=self.device)
            if n_valid:
                new[:n_valid].copy_(old[:n_valid])
            setattr(self, name, new)
        self.cap = new_cap
        self._struct = None
        self._ptrs = None


class CountBound:
    """Host upper bound of a device particle count, kept without per-step
    synchronisation.  `grow(n)` adds possible new particles; `snapshot()`
    returns the cumulative growth to store with an asynchronous copy of the
    count; `refresh(count, snap)` tightens the bound with a count observed
    when the growth was `snap` (a step never adds particles except through
    `grow`, so count + growth since then is a valid bound)."""

    def __init__(self):
        self.bound = 0
        self.grown = 0
        self.floor = 0

    def grow(self, n):
        self.bound += int(n)
        self.grown += int(n)

    def top_up(self, target):
        """A launch that raises the count to at least `target` (population
        hold): count' = max(count, target) <= max(bound, target).  Later
        refreshes bound the count by max(observed, target) + growth since,
        since every hold since the observation raised it to target at most."""
        self.floor = max(self.floor, int(target))
        self.bound = max(self.bound, int(target))

    def snapshot(self):
        return self.grown

    def refresh(self, count, snap):
        self.bound = min(self.bound, max(int(count), self.floor) + (self.grown - snap))

    def reset(self, exact):
        self.bound = int(exact)
