"""ctypes wrapper of the CPU oracle (oracle/build/libsphmhd_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker of the HIP path.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "libsphmhd_oracle.so"
# the same oracle over an operation-counting scalar (flopcount.cpp)
COUNTED_LIB = HERE / "build" / "libsphmhd_oracle_counted.so"
# the same oracle over a taped reverse-mode scalar (tape.cpp)
TAPED_LIB = HERE / "build" / "libsphmhd_oracle_taped.so"
_libs: dict = {}


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib(path: Path = LIB):
    if path not in _libs:
        if not path.exists():
            build()
        L = ctypes.CDLL(str(path))
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_int64)]
        L.orc_destroy.argtypes = [vp]
        L.orc_prop.restype = dp
        L.orc_prop.argtypes = [vp, ctypes.c_char_p]
        L.orc_itype.restype = ctypes.POINTER(ctypes.c_int32)
        L.orc_itype.argtypes = [vp]
        L.orc_prop_name.restype = ctypes.c_char_p
        L.orc_prop_name.argtypes = [ctypes.c_int]
        for name in ("orc_eval", "orc_initialize", "orc_stage1", "orc_stage2", "orc_advance_clock",
                     "orc_sync_initial_state", "orc_warm_start"):
            getattr(L, name).argtypes = [vp]
        L.orc_step.argtypes = [vp, i64]
        L.orc_steps_taken.restype = i64
        L.orc_steps_taken.argtypes = [vp]
        L.orc_get_clock.argtypes = [vp, dp, dp]
        L.orc_set_clock.argtypes = [vp, dp, dp]
        L.orc_get_pending_dt.argtypes = [vp, dp]
        L.orc_totals.argtypes = [vp, dp]
        L.orc_last_iterations.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
        L.orc_cells.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_int32)]
        for name in ("orc_neighbours", "orc_neighbours_brute"):
            getattr(L, name).argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32), i64,
                                         ctypes.POINTER(ctypes.c_int64)]
        L.orc_call.argtypes = [vp, ctypes.c_int, i64, i64, dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        for name in ("orc_kernel_w", "orc_kernel_dwdq", "orc_kernel_gradh"):
            getattr(L, name).restype = ctypes.c_double
            getattr(L, name).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_kernel_grad.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double, dp]
        L.orc_kernel_radius.restype = ctypes.c_double
        L.orc_kernel_radius.argtypes = [ctypes.c_int]
        if path == COUNTED_LIB:
            L.orcflop_get.argtypes = [ctypes.POINTER(ctypes.c_int64)] * 3
        if path == TAPED_LIB:
            pp = ctypes.POINTER(dp)
            L.tape_rollout.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_int64), pp, ctypes.POINTER(ctypes.c_int32),
                                       i64, i32, dp, dp, pp, dp, ctypes.POINTER(ctypes.c_int64)]
        _libs[path] = L
    return _libs[path]


def prop_names() -> list[str]:
    L = lib()
    return [L.orc_prop_name(k).decode() for k in range(L.orc_nprops())]


# body identifiers (sphmhd_oracle.h)
BODIES = ["DENSITY_INIT", "DENSITY_LOOP", "DENSITY_POST", "EOS", "RHO_FROM_H", "AH_FROM_RHO", "AH_FROM_DIV",
          "MAGSTRESS_LOOP", "MHDACC_INIT", "MHDACC_LOOP", "MHDACC_POST", "CD_INIT", "CD_LOOP", "CD_POST",
          "STRAIN", "BRECON", "MAGRATES_INIT", "MAGRATES_LOOP", "MAGRATES_POST", "ADAPTIVE_DT",
          "ENERGY_LIMITER", "DIVB_CORRECTION", "STEP_INITIALIZE", "STEP_STAGE1", "STEP_STAGE2"]
BODY = {name: k for k, name in enumerate(BODIES)}


class OracleSystem:
    """CPU oracle over the same instances/offsets as a DeviceSystem."""

    def flop_counts(self):
        """(body, sym, pairs) since the last flop_reset (an
        OracleSystem(counted=True)): per pass (density, Cullen-Dehnen, MHD
        accelerations, magnetic stress, magnetic state rates) the counted
        operations of the loop bodies and the pairs visited, and per pair
        symbol (r, rho_ij, W_i, grad W_i, grad W_j, dW_i/dh) its operations
        summed over every pass."""
        body, sym, pairs = (ctypes.c_int64 * 5)(), (ctypes.c_int64 * 6)(), (ctypes.c_int64 * 5)()
        self.L.orcflop_get(body, sym, pairs)
        return tuple(np.array(list(a), dtype=np.int64) for a in (body, sym, pairs))

    def flop_reset(self):
        self.L.orcflop_reset()

    def __init__(self, params, offsets=None, n=None, threads: int | None = None, counted: bool = False):
        from sphmhd._abi import Params  # noqa: F401  (layout shared with the library)
        self.L = lib(COUNTED_LIB if counted else LIB)
        plist = [params] if not isinstance(params, (list, tuple)) else list(params)
        self.params = plist
        self.params_arr = (type(plist[0]) * len(plist))(*plist)
        if offsets is None:
            offsets = [0, int(n)]
        self.offsets = np.asarray(offsets, dtype=np.int64)
        self.ninst = len(plist)
        self.n = int(self.offsets[-1])
        if threads is not None:
            self.L.orc_set_threads(int(threads))
        self.h = self.L.orc_create(ctypes.cast(self.params_arr, ctypes.c_void_p), self.ninst,
                                   self.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        if not self.h:
            raise ValueError("orc_create rejected the parameters")

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prop(self, name: str) -> np.ndarray:
        p = self.L.orc_prop(self.h, name.encode())
        if not p:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=(self.n,))

    def itype(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.L.orc_itype(self.h), shape=(self.n,))

    def set_fields(self, values):
        for name, val in values.items():
            if name == "itype":
                self.itype()[:] = np.asarray(val, dtype=np.int32)
            else:
                self.prop(name)[:] = np.asarray(val, dtype=np.float64)

    def get(self, name):
        return self.itype().copy() if name == "itype" else self.prop(name).copy()

    def state(self, names):
        return {n: self.get(n) for n in names}

    def eval(self):
        return self.L.orc_eval(self.h)

    def initialize(self):
        self.L.orc_initialize(self.h)

    def stage1(self):
        self.L.orc_stage1(self.h)

    def stage2(self):
        self.L.orc_stage2(self.h)

    def advance_clock(self):
        self.L.orc_advance_clock(self.h)

    def step(self, nsteps=1):
        return self.L.orc_step(self.h, int(nsteps))

    def steps_taken(self) -> int:
        return int(self.L.orc_steps_taken(self.h))

    def sync_initial_state(self):
        self.L.orc_sync_initial_state(self.h)

    def warm_start(self):
        rc = self.L.orc_warm_start(self.h)
        if rc:
            raise RuntimeError(f"oracle warm start failed: {rc}")

    def clock(self):
        t = np.zeros(self.ninst)
        dt = np.zeros(self.ninst)
        self.L.orc_get_clock(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             dt.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return t, dt

    def set_clock(self, t, dt):
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (self.ninst,)))
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.ninst,)))
        self.L.orc_set_clock(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             dt.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def pending_dt(self):
        dt = np.zeros(self.ninst)
        self.L.orc_get_pending_dt(self.h, dt.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return dt

    def last_iterations(self):
        it = np.zeros(self.ninst, dtype=np.int32)
        self.L.orc_last_iterations(self.h, it.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return it

    def totals(self):
        out = np.zeros((self.ninst, 9))
        self.L.orc_totals(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def cells(self):
        keys = np.zeros(self.n, dtype=np.uint32)
        perm = np.zeros(self.n, dtype=np.int32)
        self.L.orc_cells(self.h, keys.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                         perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return keys, perm

    def _nb(self, fn):
        offs = np.zeros(self.n + 1, dtype=np.int64)
        total = ctypes.c_int64(0)
        cap = 64 * self.n + 64
        for _ in range(2):
            idx = np.zeros(cap, dtype=np.int32)
            rc = fn(self.h, offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                    idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cap, ctypes.byref(total))
            if rc == 0:
                return offs, idx[: total.value]
            cap = int(total.value)
        raise RuntimeError("neighbour export failed")

    def neighbours(self):
        return self._nb(self.L.orc_neighbours)

    def neighbours_brute(self):
        return self._nb(self.L.orc_neighbours_brute)

    def call(self, body: str, d: int, s: int = 0, sym=None) -> int:
        v = np.zeros(15) if sym is None else np.asarray(sym, dtype=np.float64)
        return self.L.orc_call(self.h, BODY[body], d, s, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))


def kernel_w(kernel: int, dim: int, r: float, h: float) -> float:
    return lib().orc_kernel_w(kernel, dim, r, h)


def kernel_dwdq(kernel: int, dim: int, r: float, h: float) -> float:
    return lib().orc_kernel_dwdq(kernel, dim, r, h)


def kernel_gradh(kernel: int, dim: int, r: float, h: float) -> float:
    return lib().orc_kernel_gradh(kernel, dim, r, h)


def kernel_grad(kernel: int, dim: int, xij, r: float, h: float) -> np.ndarray:
    x = np.asarray(xij, dtype=np.float64)
    g = np.zeros(3)
    lib().orc_kernel_grad(kernel, dim, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), r, h,
                          g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return g


def kernel_radius(kernel: int) -> float:
    return lib().orc_kernel_radius(kernel)


def tape_rollout(params, offsets, fields, itype=None, nsteps: int = 0, warm: bool = True, weights=None):
    """Reference gradient from the taped oracle (tape.cpp): the objective
    sum_q weights[q] * total_q per instance after the warm start (``warm``)
    and ``nsteps`` PEC steps from the initial state ``fields`` (property name
    -> array; the rest zero), and the gradient of the objectives' sum with
    respect to every initial property and to the parameters of TPARAMS.

    Returns (values [ninst], {property: gradient array}, grad_params
    [ninst, 6], tape nodes)."""
    L = lib(TAPED_LIB)
    names = prop_names()
    plist = [params] if not isinstance(params, (list, tuple)) else list(params)
    parr = (type(plist[0]) * len(plist))(*plist)
    offsets = np.asarray(offsets, dtype=np.int64)
    ninst, n = len(plist), int(offsets[-1])
    keep = []
    fptr = (ctypes.POINTER(ctypes.c_double) * len(names))()
    gptr = (ctypes.POINTER(ctypes.c_double) * len(names))()
    grads = {}
    for q, nm in enumerate(names):
        if nm in fields:
            a = np.ascontiguousarray(np.asarray(fields[nm], dtype=np.float64))
            keep.append(a)
            fptr[q] = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        g = np.zeros(n)
        grads[nm] = g
        gptr[q] = g.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    it = None
    if itype is not None:
        it = np.ascontiguousarray(np.asarray(itype, dtype=np.int32))
    elif "itype" in fields:
        it = np.ascontiguousarray(np.asarray(fields["itype"], dtype=np.int32))
    else:
        it = np.ones(n, dtype=np.int32)
    w = np.zeros(9)
    if weights is None:
        w[4] = w[6] = 1.0  # E_kin + E_mag
    else:
        w[:] = np.asarray(weights, dtype=np.float64)
    vals = np.zeros(ninst)
    gpar = np.zeros((ninst, 6))
    nodes = ctypes.c_int64(0)
    rc = L.tape_rollout(ctypes.cast(parr, ctypes.c_void_p), ninst,
                        offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), fptr,
                        it.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), int(nsteps), 1 if warm else 0,
                        w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                        vals.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), gptr,
                        gpar.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(nodes))
    if rc:
        raise RuntimeError(f"taped oracle rollout failed: {rc}")
    return vals, grads, gpar, int(nodes.value)
