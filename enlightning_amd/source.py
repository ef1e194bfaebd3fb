"""Source injection on the device: enl_add_source (enlightning::addSource,
enlightning.C:2140-2307) and the lightning-channel file reader
enl_read_lightning (enlightning.C:2333-2389).

Defaults follow the reference input (input.txt:20-38)."""
import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from .patch import _darr, _ptr, use_torch_stream

POINT, SINE, LIGHTNING = 10, 20, 30  # enlightning.C:55-57
PULSE, DRIVEN = 10, 20               # enlightning.C:60-61


@dataclass
class Source:
    mode: int = PULSE
    type: int = POINT
    amp: float = 5.0e5
    spread: float = 0.01
    pos: tuple = (0.2, 0.2)
    sine_amp: float = 1.0
    sine_freq: float = 1.0
    omega: float = 1000.0
    coords: Optional[np.ndarray] = None  # lightning: (nseg, 4) x1 y1 x2 y2

    def to_struct(self):
        s = _lib.EnlSource()
        s.mode, s.type = int(self.mode), int(self.type)
        s.amp, s.spread = float(self.amp), float(self.spread)
        s.pos[0], s.pos[1] = float(self.pos[0]), float(self.pos[1])
        s.sine_amp, s.sine_freq, s.omega = float(self.sine_amp), float(self.sine_freq), float(self.omega)
        if self.coords is not None:
            c = np.ascontiguousarray(self.coords, dtype=np.float64).reshape(-1, 4)
            s.nseg = c.shape[0]
            s.coords = c.ctypes.data
            s._keep = c  # the array must outlive the call
        else:
            s.nseg, s.coords = 0, None
        return s


def should_add_source(src, iteration, pulse_times):
    """enlightning::shouldAddSource (enlightning.C:2118-2135), evaluated
    before the main loop increments the iteration number: a DRIVEN source is
    drawn on every iteration from its first pulse time on, a PULSE source on
    the iterations with iteration + 1 equal to one of its pulse times."""
    if src.mode == DRIVEN:
        return iteration >= pulse_times[0]
    if src.mode == PULSE:
        return any(iteration + 1 == t for t in pulse_times)
    return False


def add_source(box, patch_xlo, dx, src, iteration, pulse_time0, dt, source, sy=None):
    """Writes the patch's mass source into the device array `source` (interior
    cells, row stride sy; default the patch width)."""
    lib = _lib.load()
    use_torch_stream()
    st = src.to_struct()
    nx = box[1] - box[0] + 1
    # a view into a larger array is fine: the rows are addressed through sy
    ptr = ctypes.c_void_p(source.data_ptr()) if hasattr(source, "data_ptr") else _ptr(source)
    rc = lib.enl_add_source(box[0], box[1], box[2], box[3], _darr(patch_xlo), _darr(dx), ctypes.byref(st),
                            int(iteration), int(pulse_time0), float(dt), ptr, int(sy if sy else nx))
    _lib.check(rc)


def read_lightning(path, max_segments=100000):
    """Segments of a lightning-channel file, as an (n, 4) array."""
    lib = _lib.load()
    buf = np.zeros((max_segments, 4))
    n = lib.enl_read_lightning(str(path).encode(), buf.ctypes.data, max_segments)
    _lib.check(n if n < 0 else 0)
    return buf[:min(n, max_segments)].copy()


def sample_mics(box, patch_xlo, patch_xhi, dx, mic_x, mic_y, pres, sy=None):
    """Pressure at the microphones inside the patch (enl_sample_mics,
    enlightning.C:2414-2436) from the device pressure `pres` (interior cells,
    row stride sy; a view into a ghosted array is fine): device tensors
    (values, hit)."""
    import torch
    lib = _lib.load()
    use_torch_stream()
    mx = torch.as_tensor(np.asarray(mic_x, dtype=np.float64), device=pres.device)
    my = torch.as_tensor(np.asarray(mic_y, dtype=np.float64), device=pres.device)
    out = torch.zeros(len(mx), dtype=torch.float64, device=pres.device)
    hit = torch.zeros(len(mx), dtype=torch.int32, device=pres.device)
    nx = box[1] - box[0] + 1
    rc = lib.enl_sample_mics(box[0], box[1], box[2], box[3], _darr(patch_xlo), _darr(patch_xhi), _darr(dx), len(mx),
                             ctypes.c_void_p(mx.data_ptr()), ctypes.c_void_p(my.data_ptr()),
                             ctypes.c_void_p(pres.data_ptr()), int(sy if sy else nx), ctypes.c_void_p(out.data_ptr()),
                             ctypes.c_void_p(hit.data_ptr()))
    _lib.check(rc)
    return out, hit
