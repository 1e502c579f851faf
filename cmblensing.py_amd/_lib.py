"""ctypes binding of libcmbl.so (the C ABI in include/cmbl.h).

The library is the only compute path: if it is missing, or no HIP device is
visible, every entry point raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading
from collections import OrderedDict
from pathlib import Path

import torch

PKG = Path(__file__).resolve().parent
# CMBL_LIBRARY may point at another build of the same library (A/B builds)
LIB_PATH = Path(os.environ.get("CMBL_LIBRARY", str(PKG / "libcmbl.so")))

CMBL_F32, CMBL_F64 = 0, 1
CMBL_PHI_MAP, CMBL_PHI_FOURIER = 0, 1

_lib = None
_lock = threading.Lock()

c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
_SIGS = {
    "cmbl_version": (ctypes.c_char_p, []),
    "cmbl_last_error": (ctypes.c_char_p, []),
    "cmbl_plan_create": (c_int, [ctypes.POINTER(c_void_p), c_int, c_double, c_int, c_int]),
    "cmbl_plan_destroy": (c_int, [c_void_p]),
    "cmbl_phi_prepare": (c_int, [c_void_p, c_void_p, c_int, c_int, ctypes.POINTER(c_void_p),
                                 ctypes.POINTER(c_double), c_void_p]),
    "cmbl_phi_release": (c_int, [c_void_p]),
    "cmbl_phi_release_async": (c_int, [c_void_p, c_void_p]),
    "cmbl_phi_min_det": (c_int, [c_void_p, ctypes.POINTER(c_double)]),
    "cmbl_phi_prepare_batch": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int,
                                       ctypes.POINTER(c_void_p), ctypes.POINTER(c_double),
                                       c_void_p]),
    "cmbl_phi_derivatives": (c_void_p, [c_void_p]),
    "cmbl_lenseflow": (c_int, [c_void_p, ctypes.POINTER(c_void_p), c_void_p, c_void_p, c_int,
                               c_int, c_int, c_int, c_int, c_void_p]),
    "cmbl_lenseflow_noalloc": (c_int, [c_void_p, ctypes.POINTER(c_void_p), c_void_p, c_void_p,
                                       c_int, c_int, c_int, c_int, c_int, c_void_p]),
    "cmbl_rfft2": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "cmbl_irfft2": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "cmbl_grf_fourier": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                 ctypes.c_ulonglong, ctypes.c_ulonglong, c_void_p]),
    "cmbl_gradf_logpdf": (c_int, [c_void_p, ctypes.POINTER(c_void_p), c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_int, c_int, c_int, c_void_p]),
    "cmbl_lensed_cotangent": (c_int, [c_void_p, ctypes.POINTER(c_void_p), c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                                      c_void_p]),
    "cmbl_gradf_final": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                 c_void_p]),
    "cmbl_dot_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p,
                             c_void_p]),
    "cmbl_cg_update_xrz": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "cmbl_cg_update_p": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                 c_void_p]),
    "cmbl_cg_step": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                             c_void_p]),
    "cmbl_lenseflow_vjp": (c_int, [c_void_p, ctypes.POINTER(c_void_p), c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p]),
    "cmbl_exec_create": (c_int, [ctypes.POINTER(c_void_p), c_void_p, c_int, c_int, c_int, c_int,
                                 c_int]),
    "cmbl_exec_run": (c_int, [c_void_p, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                              ctypes.POINTER(c_void_p), c_int, c_void_p]),
    "cmbl_exec_destroy": (c_int, [c_void_p]),
    "cmbl_plan_set_timing": (c_int, [c_void_p, c_int]),
    "cmbl_plan_get_timing": (c_int, [c_void_p, ctypes.POINTER(c_double), c_int]),
}


class CmblError(RuntimeError):
    pass


def exported_symbols():
    return list(_SIGS)


def load(require_device: bool = True):
    """Load libcmbl.so once.  Raises if the library is not built or, with
    require_device, if no HIP device is visible."""
    global _lib
    if require_device and not torch.cuda.is_available():
        raise CmblError("no HIP device visible: the LenseFlow path runs only on the GPU")
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise CmblError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
            lib = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in _SIGS.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(status: int, what: str):
    if status != 0:
        msg = load(False).cmbl_last_error().decode()
        raise CmblError(f"{what} failed (status {status}): {msg}")


def stream_ptr(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float32:
        return CMBL_F32
    if dtype == torch.float64:
        return CMBL_F64
    raise CmblError(f"unsupported dtype {dtype} (float32 or float64)")


class Plan:
    """Owns a cmbl_plan: geometry, dtype and stage scratch for max_maps maps."""

    def __init__(self, N: int, theta_pix: float, dtype: torch.dtype, max_maps: int, device):
        self.lib = load()
        self.N, self.theta_pix, self.dtype, self.max_maps = N, float(theta_pix), dtype, max_maps
        self.device = torch.device(device)
        h = c_void_p()
        with torch.cuda.device(self.device):
            check(self.lib.cmbl_plan_create(ctypes.byref(h), N, self.theta_pix, dtype_code(dtype),
                                            max_maps), "cmbl_plan_create")
        self.handle = h

    def set_timing(self, enable: bool):
        check(self.lib.cmbl_plan_set_timing(self.handle, int(bool(enable))), "set_timing")

    def get_timing(self, reset: bool = True) -> dict:
        out = (c_double * 6)()
        check(self.lib.cmbl_plan_get_timing(self.handle, out, int(bool(reset))), "get_timing")
        return {"col_ms": out[0], "col_n": int(out[1]), "row_ms": out[2], "row_n": int(out[3]),
                "other_ms": out[4], "other_n": int(out[5])}

    def close(self):
        if self.handle:
            torch.cuda.synchronize(self.device)
            self.lib.cmbl_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _phi_basis(phi: torch.Tensor, N: int) -> int:
    if tuple(phi.shape[-2:]) == (N, N) and not phi.is_complex():
        return CMBL_PHI_MAP
    if tuple(phi.shape[-2:]) == (N, N // 2 + 1):
        return CMBL_PHI_FOURIER
    raise CmblError(f"phi shape {tuple(phi.shape)} does not match N={N}")


class PhiHandle:
    """A prepared phi: float64 derivatives and the p(t) table for nsteps.

    ``streams`` records the streams its tables were used on (preparation and
    every library call that read them); close() releases the handle without
    waiting when that is only the current stream, and after waiting for the
    device otherwise.  ``min_det`` (the caustic guard) is read from the device
    the first time it is asked for."""

    @classmethod
    def _wrap(cls, plan: Plan, handle: c_void_p, nsteps: int, stream: int,
              ready: "torch.cuda.Event") -> "PhiHandle":
        self = cls.__new__(cls)
        self.lib, self.plan, self.nsteps = plan.lib, plan, nsteps
        self.handle, self._min_det, self.streams = handle, None, {stream}
        self.prep_stream, self.ready = stream, ready
        return self

    def order_on_current_stream(self):
        """Make the current stream wait for this handle's preparation when it
        was prepared on another stream (a cached handle used from a second
        stream would otherwise read tables still being written)."""
        cur = torch.cuda.current_stream(self.plan.device)
        if cur.cuda_stream != self.prep_stream:
            cur.wait_event(self.ready)

    @property
    def min_det(self) -> float:
        if self._min_det is None:
            md = c_double()
            with torch.cuda.device(self.plan.device):
                check(self.lib.cmbl_phi_min_det(self.handle, ctypes.byref(md)), "cmbl_phi_min_det")
            self._min_det = md.value
        return self._min_det

    def __init__(self, plan: Plan, phi: torch.Tensor, nsteps: int, with_min_det: bool = False):
        """Prepare phi on the current stream.  with_min_det: also read the
        caustic guard now (waits for this stream; the handle is then complete
        on return), otherwise nothing waits."""
        self.lib = plan.lib
        self.plan = plan
        self.nsteps = nsteps
        N = plan.N
        phi = phi.to(device=plan.device, dtype=torch.complex128 if phi.is_complex() else torch.float64)
        phi = phi.contiguous()
        if phi.shape == (N, N):
            basis = CMBL_PHI_MAP
        elif phi.shape == (N, N // 2 + 1):
            basis = CMBL_PHI_FOURIER
        else:
            raise CmblError(f"phi shape {tuple(phi.shape)} does not match N={N}")
        h = c_void_p()
        md = c_double()
        st = stream_ptr(plan.device)
        with torch.cuda.device(plan.device):
            check(self.lib.cmbl_phi_prepare(plan.handle, phi.data_ptr(), basis, nsteps,
                                            ctypes.byref(h),
                                            ctypes.byref(md) if with_min_det else None, st),
                  "cmbl_phi_prepare")
        self.handle = h
        self._min_det = md.value if with_min_det else None
        self.streams = {st}
        self.prep_stream = st
        self.ready = _record_event(plan.device)

    def derivatives(self) -> torch.Tensor:
        """The five float64 phi derivative maps as a [5, N, N] tensor (a copy)."""
        N = self.plan.N
        ptr = self.lib.cmbl_phi_derivatives(self.handle)
        out = torch.empty((5, N, N), dtype=torch.float64, device=self.plan.device)
        torch.cuda.current_stream(self.plan.device).synchronize()
        _memcpy_d2d(out.data_ptr(), ptr, out.numel() * 8)
        return out

    def close(self):
        if self.handle:
            cur = stream_ptr(self.plan.device)
            with torch.cuda.device(self.plan.device):
                if self.streams <= {cur}:
                    self.lib.cmbl_phi_release_async(self.handle, cur)
                else:
                    torch.cuda.synchronize(self.plan.device)
                    self.lib.cmbl_phi_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def note_streams(phis, stream: int):
    """Record that the tables of these handles are read on `stream`."""
    for h in phis:
        h.streams.add(stream)


class Executor:
    """A cmbl_exec: a batched LenseFlow of fixed capacity (max_batch requests of
    ncomp components) captured once into a hipGraph on the plan's staging
    buffer.  The executor keeps its plan alive; runs of executors sharing a
    plan must be issued on one stream."""

    def __init__(self, plan: Plan, max_batch: int, ncomp: int, nsteps: int, direction: int,
                 adjoint: bool):
        self.lib, self.plan = plan.lib, plan
        self.max_batch, self.ncomp, self.nsteps = max_batch, ncomp, nsteps
        self.direction, self.adjoint = int(direction), bool(adjoint)
        h = c_void_p()
        with torch.cuda.device(plan.device):
            check(self.lib.cmbl_exec_create(ctypes.byref(h), plan.handle, max_batch, ncomp, nsteps,
                                            self.direction, int(self.adjoint)), "cmbl_exec_create")
        self.handle = h
        self._arrs = ((c_void_p * max_batch)(), (c_void_p * max_batch)(), (c_void_p * max_batch)())

    def run(self, phis, in_ptrs, out_ptrs, stream: int):
        """Lens len(phis) requests: phis are PhiHandles, in_ptrs/out_ptrs device
        addresses of [ncomp, N, N] maps; enqueued on `stream` (a hipStream_t)."""
        count = len(phis)
        a, i, o = self._arrs
        note_streams(phis, stream)
        for b in range(count):
            a[b] = phis[b].handle.value
        i[:count] = in_ptrs
        o[:count] = out_ptrs
        check(self.lib.cmbl_exec_run(self.handle, a, i, o, count, stream), "cmbl_exec_run")

    def close(self):
        if self.handle:
            torch.cuda.synchronize(self.plan.device)
            self.lib.cmbl_exec_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _memcpy_d2d(dst: int, src: int, nbytes: int):
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [c_void_p, c_void_p, ctypes.c_size_t, c_int]
    rc = hip.hipMemcpy(c_void_p(dst), c_void_p(src), nbytes, 3)  # DeviceToDevice
    if rc != 0:
        raise CmblError(f"hipMemcpy failed ({rc})")


# --------------------------------------------------------------------------
# caches
# --------------------------------------------------------------------------
_plans: dict = {}
_phis: "OrderedDict" = OrderedDict()
PHI_CACHE_SIZE = 64
# the caches are shared by host threads (the serving Wiener-filter lane runs on
# one of its own)
_cache_lock = threading.RLock()


def _locked(fn):
    import functools

    @functools.wraps(fn)
    def wrapper(*a, **k):
        with _cache_lock:
            return fn(*a, **k)
    return wrapper


@_locked
def get_plan(N: int, theta_pix: float, dtype: torch.dtype, nmaps: int, device) -> Plan:
    """The cached plan of (device, geometry, dtype) for the current stream.

    A plan's scratch (stage spectra, accumulator, phi-preparation and
    Wiener-filter temporaries) is used by every call on it, so plans are kept
    per stream: calls issued on different streams never share scratch.  When a
    larger plan is needed the old one leaves the cache but is not destroyed:
    whoever still holds it (a Wiener-filter process between its steps, say)
    keeps a working plan, which is freed with its last reference."""
    device = torch.device(device)
    if device.index is None:
        device = torch.device("cuda", torch.cuda.current_device())
    key = (device.index, N, round(float(theta_pix), 12), dtype,
           torch.cuda.current_stream(device).cuda_stream)
    plan = _plans.get(key)
    if plan is None or plan.max_maps < nmaps:
        cap = max(nmaps, 2 * plan.max_maps if plan else nmaps)
        old, plan = plan, Plan(N, theta_pix, dtype, cap, device)
        _plans[key] = plan
        # cached phi handles need no plan of their own (their memory is the
        # library's): point them at the new plan so they do not keep the old
        # plan's scratch alive
        if old is not None:
            for v in _phis.values():
                if v[2].plan is old:
                    v[2].plan = plan
    return plan


def prepare_batch(plan: Plan, phis: torch.Tensor, nsteps: int):
    """PhiHandles for every phi of a [count, N, N] (MAP) or [count, N, N//2+1]
    (FOURIER) tensor, prepared in one cmbl_phi_prepare_batch call."""
    N = plan.N
    count = int(phis.shape[0])
    basis = _phi_basis(phis, N)
    phis = phis.to(device=plan.device,
                   dtype=torch.complex128 if basis == CMBL_PHI_FOURIER else torch.float64)
    phis = phis.contiguous()
    hs = (c_void_p * count)()
    st = stream_ptr(plan.device)
    with torch.cuda.device(plan.device):
        check(plan.lib.cmbl_phi_prepare_batch(plan.handle, phis.data_ptr(), basis, nsteps, count,
                                              hs, None, st), "cmbl_phi_prepare_batch")
    ready = _record_event(plan.device)
    return [PhiHandle._wrap(plan, c_void_p(hs[i]), nsteps, st, ready) for i in range(count)]


def _record_event(device) -> "torch.cuda.Event":
    ev = torch.cuda.Event()
    ev.record(torch.cuda.current_stream(device))
    return ev


@_locked
def get_phi_batch(plan: Plan, phi: torch.Tensor, nsteps: int):
    """Prepared phis for every element of a batched phi [B, ...] (cached like
    get_phi per element; any miss prepares the whole batch in one call)."""
    import weakref
    B = int(phi.shape[0])
    reserve_phi_cache(B)
    pkey = (plan.device.index, plan.N, round(plan.theta_pix, 12), plan.dtype)
    keys = [(pkey, id(phi), i, nsteps) for i in range(B)]
    hit = []
    for k in keys:
        ent = _phis.get(k)
        if ent is None or ent[0]() is not phi or ent[1] != phi._version:
            break
        hit.append(ent[2])
    if len(hit) == B:
        for k, h in zip(keys, hit):
            _phis.move_to_end(k)
            h.order_on_current_stream()
        return hit
    handles = prepare_batch(plan, phi, nsteps)
    ref = weakref.ref(phi)
    for k, h in zip(keys, handles):
        _phis.pop(k, None)
        _phis[k] = (ref, phi._version, h)
    _trim_phi_cache()
    return handles


def reserve_phi_cache(count: int):
    """Grow the phi cache so that `count` handles fetched back to back all stay
    alive (the oldest entry is released when the cache overflows)."""
    global PHI_CACHE_SIZE
    PHI_CACHE_SIZE = max(PHI_CACHE_SIZE, int(count))


@_locked
def get_phi(plan: Plan, phi: torch.Tensor, nsteps: int, index=None) -> PhiHandle:
    """Prepared phi for (plan geometry, the tensor object and its version,
    nsteps) -- or for ``phi[index]`` of a batched phi tensor.  Entries are
    keyed by object identity (checked through a weak reference), never by
    address: torch's caching allocator hands freed addresses to new tensors."""
    import weakref
    pkey = (plan.device.index, plan.N, round(plan.theta_pix, 12), plan.dtype)
    key = (pkey, id(phi), index, nsteps)
    ent = _phis.get(key)
    if ent is not None:
        ref, version, h = ent
        if ref() is phi and version == phi._version:
            _phis.move_to_end(key)
            h.order_on_current_stream()
            return h
        del _phis[key]
    h = PhiHandle(plan, phi if index is None else phi[index], nsteps)
    _phis[key] = (weakref.ref(phi), phi._version, h)
    _trim_phi_cache()
    return h


def _trim_phi_cache():
    """Drop the least recently used entries beyond PHI_CACHE_SIZE.  A dropped
    handle is not released here: a caller still holding it (the phi handles
    of a Wiener-filter process, say) keeps using it, and the handle is
    released with its last reference (PhiHandle.__del__)."""
    while len(_phis) > PHI_CACHE_SIZE:
        _phis.popitem(last=False)


@_locked
def clear_caches():
    for v in _phis.values():
        v[2].close()
    _phis.clear()
    for p in _plans.values():
        p.close()
    _plans.clear()


def lenseflow(plan: Plan, phis, x: torch.Tensor, out: torch.Tensor, batch: int, ncomp: int,
              nsteps: int, direction: int, adjoint: bool):
    arr = (c_void_p * batch)(*[p.handle for p in phis])
    st = stream_ptr(plan.device)
    note_streams(phis, st)
    with torch.cuda.device(plan.device):
        check(plan.lib.cmbl_lenseflow(plan.handle, arr, x.data_ptr(), out.data_ptr(), batch, ncomp,
                                      nsteps, int(direction), int(bool(adjoint)), st),
              "cmbl_lenseflow")


NOALLOC_SCRATCH_BYTES = 64 << 20


def lenseflow_noalloc(plan: Plan, phis, x: torch.Tensor, out: torch.Tensor, batch: int,
                      ncomp: int, nsteps: int, direction: int, adjoint: bool):
    """cmbl_lenseflow_noalloc: cmbl_lenseflow with its stage buffers in the
    library's static per-GPU scratch (NOALLOC_SCRATCH_BYTES); the call itself
    allocates nothing.  Calls overlapping on one GPU race on that scratch."""
    arr = (c_void_p * max(batch, 1))(*[p.handle for p in phis])
    st = stream_ptr(plan.device)
    note_streams(phis, st)
    with torch.cuda.device(plan.device):
        check(plan.lib.cmbl_lenseflow_noalloc(plan.handle, arr, x.data_ptr(), out.data_ptr(),
                                              batch, ncomp, nsteps, int(direction),
                                              int(bool(adjoint)), st), "cmbl_lenseflow_noalloc")


def rfft2(plan: Plan, x: torch.Tensor, X: torch.Tensor, nmaps: int):
    with torch.cuda.device(plan.device):
        check(plan.lib.cmbl_rfft2(plan.handle, x.data_ptr(), X.data_ptr(), nmaps,
                                  stream_ptr(plan.device)), "cmbl_rfft2")


def irfft2(plan: Plan, X: torch.Tensor, x: torch.Tensor, nmaps: int):
    with torch.cuda.device(plan.device):
        check(plan.lib.cmbl_irfft2(plan.handle, X.data_ptr(), x.data_ptr(), nmaps,
                                   stream_ptr(plan.device)), "cmbl_irfft2")


def grf_fourier(plan: Plan, cov: torch.Tensor, mul, add, out: torch.Tensor, ntrials: int,
                ncomp: int, seed: int, stream0: int):
    with torch.cuda.device(plan.device):
        check(plan.lib.cmbl_grf_fourier(plan.handle, cov.data_ptr(), _ptr(mul), _ptr(add),
                                        out.data_ptr(), ntrials, ncomp, int(seed) & (2**64 - 1),
                                        int(stream0) & (2**64 - 1), stream_ptr(plan.device)),
              "cmbl_grf_fourier")


class WfOps(ctypes.Structure):
    """cmbl_wf_ops: device pointers of the diagonal operators (plan dtype)."""
    _fields_ = [("mb", c_void_p), ("bmcn", c_void_p), ("cfinv", c_void_p), ("beam", c_void_p),
                ("cninv", c_void_p), ("mask_pix", c_void_p)]


def _ptr(t):
    return None if t is None else t.data_ptr()


def gradf_logpdf(plan: Plan, phis, f, d, ops: "WfOps", out, batch: int, ncomp: int, nsteps: int):
    arr = (c_void_p * batch)(*[p.handle for p in phis])
    st = stream_ptr(plan.device)
    note_streams(phis, st)
    with torch.cuda.device(plan.device):
        check(plan.lib.cmbl_gradf_logpdf(plan.handle, arr, _ptr(f), _ptr(d), ctypes.byref(ops),
                                         out.data_ptr(), batch, ncomp, nsteps, st),
              "cmbl_gradf_logpdf")


def lensed_cotangent(plan: Plan, phis, f, d, ops: "WfOps", lensed, cot, batch: int, ncomp: int,
                     nsteps: int):
    arr = (c_void_p * batch)(*[p.handle for p in phis])
    st = stream_ptr(plan.device)
    note_streams(phis, st)
    with torch.cuda.device(plan.device):
        check(plan.lib.cmbl_lensed_cotangent(plan.handle, arr, _ptr(f), _ptr(d), ctypes.byref(ops),
                                             lensed.data_ptr(), cot.data_ptr(), batch, ncomp,
                                             nsteps, st), "cmbl_lensed_cotangent")


def gradf_final(plan: Plan, adj, f, cfinv, out, batch: int, ncomp: int):
    """cmbl_gradf_final: out = qu2eb(fourier(adj)) - cfinv .* Herm(f) (f or
    cfinv None: no prior term)."""
    with torch.cuda.device(plan.device):
        check(plan.lib.cmbl_gradf_final(plan.handle, adj.data_ptr(), _ptr(f), _ptr(cfinv),
                                        out.data_ptr(), batch, ncomp, stream_ptr(plan.device)),
              "cmbl_gradf_final")


def dot_f64(plan: Plan, a, b, nrhs: int, ncomp: int, out):
    b_f64 = int(b.dtype in (torch.complex128, torch.float64))
    with torch.cuda.device(plan.device):
        check(plan.lib.cmbl_dot_f64(plan.handle, a.data_ptr(), b.data_ptr(), b_f64, nrhs, ncomp,
                                    out.data_ptr(), stream_ptr(plan.device)), "cmbl_dot_f64")


def cg_update_xrz(plan: Plan, x, r, z, p, Ap, pinv, alpha, nrhs: int, ncomp: int, rz):
    with torch.cuda.device(plan.device):
        check(plan.lib.cmbl_cg_update_xrz(plan.handle, x.data_ptr(), r.data_ptr(), z.data_ptr(),
                                          p.data_ptr(), Ap.data_ptr(), pinv.data_ptr(),
                                          alpha.data_ptr(), nrhs, ncomp, rz.data_ptr(),
                                          stream_ptr(plan.device)), "cmbl_cg_update_xrz")


def cg_update_p(plan: Plan, p, z, beta, p_op, nrhs: int, ncomp: int):
    with torch.cuda.device(plan.device):
        check(plan.lib.cmbl_cg_update_p(plan.handle, p.data_ptr(), z.data_ptr(), beta.data_ptr(),
                                        _ptr(p_op), nrhs, ncomp, stream_ptr(plan.device)),
              "cmbl_cg_update_p")


def cg_step(plan: Plan, x, r, z, p, Ap, pinv, p_op, res, act, bb, rel, nrhs: int, ncomp: int):
    """cmbl_cg_step: one PCG iteration after the operator, scalars on the
    device (res updated in place; act a bool device tensor or None; rel and
    bb together or None)."""
    with torch.cuda.device(plan.device):
        check(plan.lib.cmbl_cg_step(plan.handle, x.data_ptr(), r.data_ptr(), z.data_ptr(),
                                    p.data_ptr(), Ap.data_ptr(), pinv.data_ptr(), _ptr(p_op),
                                    res.data_ptr(), _ptr(act), _ptr(bb), _ptr(rel), nrhs, ncomp,
                                    stream_ptr(plan.device)), "cmbl_cg_step")


def lenseflow_vjp(plan: Plan, phis, lensed, cot, d_field, d_phi, batch: int, ncomp: int,
                  nsteps: int, direction: int):
    arr = (c_void_p * batch)(*[p.handle for p in phis])
    st = stream_ptr(plan.device)
    note_streams(phis, st)
    with torch.cuda.device(plan.device):
        check(plan.lib.cmbl_lenseflow_vjp(plan.handle, arr, lensed.data_ptr(), cot.data_ptr(),
                                          d_field.data_ptr(), d_phi.data_ptr(), batch, ncomp,
                                          nsteps, int(direction), st), "cmbl_lenseflow_vjp")
