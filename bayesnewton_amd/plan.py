"""Series plan: the series -> chunk -> tile decomposition of the chunked scan.

Wraps bn_plan_create (include/bn_hip.h).  The reference handles one series per
model and sorts it in input_admin (bayesnewton/utils.py:234-265); a plan holds
any number of already-sorted series as CSR offsets and owns the device tables
and the workspace the chunked kernels need.
"""
import ctypes

import numpy as np
import torch

from . import _native as nat


def default_chunk_len(lengths, parallel=True):
    """Chunk length: the sequential filter (one chunk per series) when
    parallel=False, else a length that keeps every lane busy for a handful of
    steps while bounding each series to 128 x 128 chunks."""
    lengths = np.asarray(lengths, dtype=np.int64)
    nmax = int(lengths.max()) if lengths.size else 1
    if not parallel:
        return max(nmax, 1)
    c = 32
    while -(-nmax // c) > 128 * 128:
        c *= 2
    return c


class SeriesPlan:
    """Device-side decomposition of B sorted series (CSR offsets) into chunks.

    offsets: int64 [B+1] with offsets[0] = 0.  chunk_len: steps per lane.
    """

    def __init__(self, offsets, chunk_len=None, parallel=True, device=None):
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        if off.ndim != 1 or off.size < 2 or off[0] != 0 or np.any(np.diff(off) < 0):
            raise ValueError("offsets must be a non-decreasing int64 vector starting at 0")
        self.offsets = off
        self.lengths = np.diff(off)
        self.n_series = off.size - 1
        self.n_steps = int(off[-1])
        if chunk_len is None:
            chunk_len = default_chunk_len(self.lengths, parallel)
        self.chunk_len = int(chunk_len)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        L = nat.lib()
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            nat.check(L.bn_plan_create(off.ctypes.data_as(ctypes.c_void_p), self.n_series, self.chunk_len,
                                       ctypes.byref(h)))
        self._h = h
        self.n_slots = int(L.bn_plan_num_slots(h))
        self.n_chunks = int(L.bn_plan_num_chunks(h))
        self.n_tiles = int(L.bn_plan_num_tiles(h))
        self._ws = {}

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                nat.lib().bn_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    def workspace(self, desc):
        """Workspace tensor for a kernel descriptor (cached per state dimension)."""
        key = (desc.n_comp, desc.kind[0], desc.kind[1])
        ws = self._ws.get(key)
        if ws is None:
            nbytes = int(nat.lib().bn_workspace_bytes(self._h, ctypes.byref(desc)))
            if nbytes < 0:
                raise nat.NativeError("bn_workspace_bytes failed")
            ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=self.device)
            self._ws[key] = ws
        return ws

    # ---- layout conversion -----------------------------------------------------
    def gather(self, x, pad=0.0):
        """[N] or [N, w] (float32/float64, on the device) -> tiled [w, n_slots]."""
        if x.dim() == 1:
            x = x[:, None]
        x = x.contiguous()
        if x.shape[0] != self.n_steps:
            raise ValueError(f"expected {self.n_steps} rows, got {x.shape[0]}")
        w = x.shape[1]
        out = torch.empty((w, self.n_slots), dtype=x.dtype, device=self.device)
        L = nat.lib()
        if x.dtype == torch.float32:
            nat.check(L.bn_plan_gather_f32(self._h, nat.ptr(x), nat.ptr(out), w, float(pad), nat.stream_ptr()))
        elif x.dtype == torch.float64:
            nat.check(L.bn_plan_gather_f64(self._h, nat.ptr(x), nat.ptr(out), w, float(pad), nat.stream_ptr()))
        else:
            raise TypeError(f"gather supports float32/float64, got {x.dtype}")
        return out

    def gather_mask(self, mask):
        """bool/uint8 [N] -> tiled uint8 [n_slots] (padding = 1, i.e. missing)."""
        m = mask.reshape(-1).to(torch.uint8).contiguous()
        out = torch.empty(self.n_slots, dtype=torch.uint8, device=self.device)
        nat.check(nat.lib().bn_plan_gather_u8(self._h, nat.ptr(m), nat.ptr(out), 1, nat.stream_ptr()))
        return out

    def scatter(self, t):
        """tiled [w, n_slots] -> [N, w]."""
        if t.dim() == 1:
            t = t[None, :]
        t = t.contiguous()
        w = t.shape[0]
        out = torch.empty((self.n_steps, w), dtype=t.dtype, device=self.device)
        L = nat.lib()
        if t.dtype == torch.float32:
            nat.check(L.bn_plan_scatter_f32(self._h, nat.ptr(t), nat.ptr(out), w, nat.stream_ptr()))
        elif t.dtype == torch.float64:
            nat.check(L.bn_plan_scatter_f64(self._h, nat.ptr(t), nat.ptr(out), w, nat.stream_ptr()))
        else:
            raise TypeError(f"scatter supports float32/float64, got {t.dtype}")
        return out
