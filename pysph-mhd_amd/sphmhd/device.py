"""Device-resident particle system driven through the C ABI of libsphmhd.so.

``DeviceSystem`` owns one FP64 storage tensor with one row per bound
property (SoA, the ParticleArray layout) plus an int32 ``itype`` row, creates
a library context on torch's current HIP stream and binds every row by raw
device pointer.  PyTorch is only the allocator and stream carrier here: all
arithmetic of the step runs in the library's HIP kernels.
"""
from __future__ import annotations

import ctypes
from typing import Iterable, Mapping, Sequence

import numpy as np
import torch

from . import _abi
from ._abi import FIELD_INDEX, FIELDS, NFIELDS, NTOTALS, TOTALS, Params, Stats, Timing, check, load_library


def _as_params_array(params: Params | Sequence[Params]):
    plist = [params] if isinstance(params, Params) else list(params)
    arr = (Params * len(plist))(*plist)
    return arr, plist


class DeviceSystem:
    """One context: an ensemble of one or more instances on one GPU."""

    def __init__(self, params: Params | Sequence[Params], offsets: Sequence[int] | None = None,
                 device: str | torch.device = "cuda", n: int | None = None):
        if not torch.cuda.is_available():
            raise RuntimeError("DeviceSystem needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.lib = load_library()
        self.params_arr, self.params = _as_params_array(params)
        if offsets is None:
            if n is None:
                raise ValueError("give offsets or n")
            offsets = [0, int(n)]
        self.offsets = np.asarray(offsets, dtype=np.int64)
        if len(self.offsets) != len(self.params) + 1:
            raise ValueError("offsets must have n_instances + 1 entries")
        self.ninst = len(self.params)
        self.n = int(self.offsets[-1])
        self.dim = int(self.params[0].dim)
        self.device = torch.device(device)
        self.storage = torch.zeros((NFIELDS, self.n), dtype=torch.float64, device=self.device)
        self.itype = torch.ones(self.n, dtype=torch.int32, device=self.device)
        self.stream = torch.cuda.current_stream(self.device)
        ctx = ctypes.c_void_p()
        off_c = self.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        check(self.lib.sphmhd_create(self.params_arr, self.ninst, off_c,
                                     ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(ctx)))
        self._ctx = ctx
        ptrs = (ctypes.c_void_p * NFIELDS)(*[self.storage[k].data_ptr() for k in range(NFIELDS)])
        check(self.lib.sphmhd_bind(self._ctx, ptrs, ctypes.c_void_p(self.itype.data_ptr())))

    # -- lifetime -------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            self.lib.sphmhd_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- fields ---------------------------------------------------------------
    def field(self, name: str) -> torch.Tensor:
        """Device view of one property (storage order)."""
        return self.storage[FIELD_INDEX[name]]

    def set_fields(self, values: Mapping[str, Iterable] ) -> None:
        for name, val in values.items():
            if name == "itype":
                self.itype.copy_(torch.as_tensor(np.asarray(val, dtype=np.int32), device=self.device))
                continue
            if name not in FIELD_INDEX:
                continue
            t = torch.as_tensor(np.asarray(val, dtype=np.float64), device=self.device)
            self.storage[FIELD_INDEX[name]].copy_(t)

    def get(self, name: str) -> np.ndarray:
        torch.cuda.synchronize(self.device)
        if name == "itype":
            return self.itype.cpu().numpy()
        return self.field(name).cpu().numpy()

    def state(self, names: Iterable[str] = FIELDS) -> dict[str, np.ndarray]:
        torch.cuda.synchronize(self.device)
        host = self.storage.cpu().numpy()
        return {nm: host[FIELD_INDEX[nm]].copy() for nm in names}

    # -- the step ---------------------------------------------------------------
    def eval(self) -> int:
        return check(self.lib.sphmhd_eval(self._ctx))

    def initialize(self) -> None:
        check(self.lib.sphmhd_pec_initialize(self._ctx))

    def stage1(self) -> None:
        check(self.lib.sphmhd_pec_stage1(self._ctx))

    def stage2(self) -> None:
        check(self.lib.sphmhd_pec_stage2(self._ctx))

    def advance_clock(self) -> None:
        check(self.lib.sphmhd_advance_clock(self._ctx))

    def step(self, nsteps: int = 1) -> int:
        """Up to nsteps PEC steps (fewer when every instance's clock reaches
        its tf first); returns the library's status code."""
        return check(self.lib.sphmhd_step(self._ctx, int(nsteps)))

    def steps_taken(self) -> int:
        """PEC steps completed by this context so far."""
        return int(self.stats()["steps"])

    def sync_initial_state(self) -> None:
        check(self.lib.sphmhd_sync_initial_state(self._ctx))

    def warm_start(self) -> None:
        check(self.lib.sphmhd_warm_start(self._ctx))

    def synchronize(self) -> None:
        check(self.lib.sphmhd_synchronize(self._ctx))

    # -- clock ------------------------------------------------------------------
    def clock(self) -> tuple[np.ndarray, np.ndarray]:
        t = np.zeros(self.ninst)
        dt = np.zeros(self.ninst)
        check(self.lib.sphmhd_get_clock(self._ctx, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        dt.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return t, dt

    def set_clock(self, t, dt) -> None:
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (self.ninst,)))
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.ninst,)))
        check(self.lib.sphmhd_set_clock(self._ctx, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        dt.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def pending_dt(self) -> np.ndarray:
        dt = np.zeros(self.ninst)
        check(self.lib.sphmhd_get_pending_dt(self._ctx, dt.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return dt

    # -- diagnostics -------------------------------------------------------------
    def totals(self) -> np.ndarray:
        out = np.zeros((self.ninst, NTOTALS))
        check(self.lib.sphmhd_totals(self._ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    STRAIN = ("dvdx_xx", "dvdx_xy", "dvdx_xz", "dvdx_yx", "dvdx_yy", "dvdx_yz", "dvdx_zx", "dvdx_zy", "dvdx_zz")

    def bind_strain(self, on: bool = True) -> None:
        """StrainDiagnostics on request: with rows bound, every evaluation that
        writes the diagnostics also writes the velocity gradient dvdx_*."""
        if on:
            self.strain_rows = torch.zeros((9, self.n), dtype=torch.float64, device=self.device)
            ptrs = (ctypes.c_void_p * 9)(*[self.strain_rows[k].data_ptr() for k in range(9)])
            check(self.lib.sphmhd_bind_strain(self._ctx, ptrs))
        else:
            check(self.lib.sphmhd_bind_strain(self._ctx, None))
            self.strain_rows = None

    def strain(self) -> dict[str, np.ndarray]:
        torch.cuda.synchronize(self.device)
        host = self.strain_rows.cpu().numpy()
        return {nm: host[k].copy() for k, nm in enumerate(self.STRAIN)}

    def totals_noalloc(self, out: torch.Tensor | None = None) -> torch.Tensor:
        """The conserved totals through sphmhd_totals_noalloc: enqueued on the
        context's stream into a device tensor [ninst, NTOTALS], with no
        allocation in the library and no synchronisation."""
        if out is None:
            out = torch.empty((self.ninst, NTOTALS), dtype=torch.float64, device=self.device)
        if out.dtype != torch.float64 or not out.is_contiguous() or out.numel() < self.ninst * NTOTALS:
            raise ValueError("out must be a contiguous float64 device tensor of ninst x NTOTALS")
        check(self.lib.sphmhd_totals_noalloc(self._ctx, ctypes.c_void_p(out.data_ptr())))
        return out

    def totals_dict(self, k: int = 0) -> dict[str, float]:
        return dict(zip(TOTALS, self.totals()[k].tolist()))

    def cells(self) -> tuple[np.ndarray, np.ndarray]:
        keys = np.zeros(self.n, dtype=np.uint32)
        perm = np.zeros(self.n, dtype=np.int32)
        check(self.lib.sphmhd_cells(self._ctx, keys.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                    perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return keys, perm

    def neighbours(self) -> tuple[np.ndarray, np.ndarray]:
        offs = np.zeros(self.n + 1, dtype=np.int64)
        total = ctypes.c_int64(0)
        cap = 0
        idx = np.zeros(1, dtype=np.int32)
        for _ in range(2):
            rc = self.lib.sphmhd_neighbours(self._ctx, offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                            idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cap,
                                            ctypes.byref(total))
            if rc == _abi.ECAP and total.value > cap:
                cap = int(total.value)
                idx = np.zeros(max(cap, 1), dtype=np.int32)
                continue
            check(rc)
            break
        return offs, idx[: int(total.value)]

    def stats(self) -> dict[str, int]:
        s = Stats()
        check(self.lib.sphmhd_get_stats(self._ctx, ctypes.byref(s)))
        return {name: getattr(s, name) for name, _ in Stats._fields_}

    def set_launch(self, blocks=None) -> None:
        """Workgroup sizes (density, lists, force), each 64, 128 or 256; None
        tunes them again in the next step call (the default: the first steps
        time every shape and keep the fastest; results never depend on it)."""
        if blocks is None:
            check(self.lib.sphmhd_set_launch(self._ctx, None))
        else:
            arr = (ctypes.c_int32 * 3)(*[int(b) for b in blocks])
            check(self.lib.sphmhd_set_launch(self._ctx, arr))

    def set_persistent(self, modes=None) -> None:
        """Launch modes (density, lists, force), each 0 (one workgroup per
        tile) or 1 (persistent workgroups pulling tiles from work queues);
        None lets the tuner choose them (results never depend on it)."""
        if modes is None:
            check(self.lib.sphmhd_set_persistent(self._ctx, None))
        else:
            arr = (ctypes.c_int32 * 3)(*[int(m) for m in modes])
            check(self.lib.sphmhd_set_persistent(self._ctx, arr))

    def launch(self) -> tuple[tuple[int, int, int], bool]:
        arr = (ctypes.c_int32 * 3)()
        tuned = ctypes.c_int32(0)
        check(self.lib.sphmhd_get_launch(self._ctx, arr, ctypes.byref(tuned)))
        return (arr[0], arr[1], arr[2]), bool(tuned.value)

    def set_timing(self, enable: bool = True) -> None:
        """Enable HIP-event timing of the hot kernels (resets the sums)."""
        check(self.lib.sphmhd_set_timing(self._ctx, 1 if enable else 0))

    def timing(self) -> dict[str, float]:
        t = Timing()
        check(self.lib.sphmhd_get_timing(self._ctx, ctypes.byref(t)))
        return {name: getattr(t, name) for name, _ in Timing._fields_}
