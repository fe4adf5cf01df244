"""Python mirror of the reference's operator API over the C ABI.

Class and method names, argument meaning and error behaviour follow the
reference's Rust backends so callers read the same:
  NttBackend::{name, forward_ntt, inverse_ntt, pointwise_mul}  (src/ntt/mod.rs:26-49)
  <Backend>::{forward_ntt_gpu, inverse_ntt_gpu}(&[u32], log_n) -> (Vec<u32>, ns)
  BbBatchNtt::{forward,inverse}_ntt_batch_gpu(&[u32], log_n, batch)  (src/ntt/babybear/batch.rs:76-180)
  CosetLdeBatch::execute(&[u32], log_n, batch, added_bits) -> LdeResult  (src/lde/mod.rs:43-107)
  NttError variants (src/ntt/mod.rs:51-69)
The reference's Metal type names are kept as aliases (MetalCtGsR2, ...).
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np

from . import _lib as L


# ------------------------------------------------------------------ errors
class NttError(Exception):
    """Mirrors the reference's NttError enum."""
    kind = "NttError"


class InvalidSize(NttError):
    kind = "InvalidSize"


class DeviceNotFound(NttError):
    kind = "DeviceNotFound"


class BufferAllocFailed(NttError):
    kind = "BufferAllocFailed"


class ShaderCompileError(NttError):
    kind = "ShaderCompileError"


class GpuExecutionError(NttError):
    kind = "GpuExecutionError"


class CorrectnessMismatch(NttError):
    kind = "CorrectnessMismatch"


class Unsupported(NttError):
    kind = "Unsupported"


class InvalidArgument(NttError):
    kind = "InvalidArgument"


class Busy(NttError):
    kind = "Busy"


_ERR = {L.ERR_INVALID_SIZE: InvalidSize, L.ERR_DEVICE_NOT_FOUND: DeviceNotFound,
        L.ERR_BUFFER_ALLOC_FAILED: BufferAllocFailed, L.ERR_KERNEL_LOAD: ShaderCompileError,
        L.ERR_GPU_EXECUTION: GpuExecutionError, L.ERR_CORRECTNESS: CorrectnessMismatch,
        L.ERR_UNSUPPORTED: Unsupported, L.ERR_INVALID_ARGUMENT: InvalidArgument,
        L.ERR_BUSY: Busy}


def _check(status: int, ctx_ptr=None):
    if status == L.OK:
        return
    lib = L.load()
    msg = lib.sfn_status_string(status).decode()
    if ctx_ptr is not None:
        detail = lib.sfn_last_error(ctx_ptr)
        if detail:
            msg = f"{msg}: {detail.decode()}"
    raise _ERR.get(status, NttError)(msg)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _p(a: np.ndarray):
    return a.ctypes.data_as(L.u32p)


# ----------------------------------------------------------------- context
class HipContext:
    """Device context (the reference's MetalContext, src/gpu.rs:26-417)."""

    def __init__(self, device: int = 0):
        self._lib = L.load()
        ptr = ctypes.c_void_p()
        _check(self._lib.sfn_ctx_create(device, ctypes.byref(ptr)))
        self.ptr = ptr
        self.lock = threading.Lock()

    def close(self):
        if getattr(self, "ptr", None):
            self._lib.sfn_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self) -> dict:
        info = L.DeviceInfo()
        _check(self._lib.sfn_device_info_get(self.ptr, ctypes.byref(info)), self.ptr)
        return {"name": info.name.decode(), "gcn_arch": info.gcn_arch.decode(),
                "compute_units": info.compute_units, "lds_bytes_per_cu": info.lds_bytes_per_cu,
                "hbm_bytes": info.hbm_bytes, "clock_khz": info.clock_khz,
                "mem_clock_khz": info.mem_clock_khz, "mem_bus_width": info.mem_bus_width}

    def check(self, status):
        _check(status, self.ptr)

    # ---- raw entry points (host pointers) ----
    def ntt(self, field, variant, direction, log_n, batch, data, layout=L.LAYOUT_SOA):
        a = _u32(data)
        out = np.empty_like(a)
        ns = ctypes.c_uint64(0)
        with self.lock:
            self.check(self._lib.sfn_ntt(self.ptr, field, variant, direction, log_n, batch, layout,
                                         _p(a), _p(out), ctypes.byref(ns)))
        return out, ns.value

    def coset_lde(self, data, log_n, batch, added_bits, shift_raw=0, layout=L.LAYOUT_SOA, natural=False):
        a = _u32(data)
        out = np.empty(batch << (log_n + added_bits), dtype=np.uint32)
        ns = ctypes.c_uint64(0)
        fn = self._lib.sfn_coset_lde_natural if natural else self._lib.sfn_coset_lde_batch
        with self.lock:
            self.check(fn(self.ptr, log_n, batch, added_bits, shift_raw, layout, _p(a), _p(out), ctypes.byref(ns)))
        return out, ns.value

    def pointwise_mul(self, field, a, b):
        a, b = _u32(a), _u32(b)
        if a.size != b.size:
            raise InvalidSize(a.size)
        out = np.empty_like(a)
        with self.lock:
            self.check(self._lib.sfn_pointwise_mul(self.ptr, field, _p(a), _p(b), _p(out), a.size))
        return out

    def twiddles_prepare(self, field, log_n):
        self.check(self._lib.sfn_twiddles_prepare(self.ptr, field, log_n))

    def twiddles_bytes(self) -> int:
        b = ctypes.c_size_t()
        self.check(self._lib.sfn_twiddles_bytes(self.ptr, ctypes.byref(b)))
        return b.value

    def twiddles_limit(self, max_bytes: int):
        """Byte budget of the per-shape tables (0 = the default)."""
        self.check(self._lib.sfn_twiddles_limit(self.ptr, max_bytes))


_default_ctx = None
_default_lock = threading.Lock()


def default_context() -> HipContext:
    global _default_ctx
    with _default_lock:
        if _default_ctx is None:
            _default_ctx = HipContext(0)
        return _default_ctx


class DeviceBuffer:
    """Device-resident u32 buffer (sfn_buf_alloc)."""

    def __init__(self, ctx: HipContext, n_u32: int):
        self.ctx, self.n = ctx, int(n_u32)
        p = ctypes.c_void_p()
        ctx.check(ctx._lib.sfn_buf_alloc(ctx.ptr, self.n * 4, ctypes.byref(p)))
        self.ptr = p

    def upload(self, host):
        a = _u32(host)
        assert a.size <= self.n
        self.ctx.check(self.ctx._lib.sfn_upload(self.ctx.ptr, self.ptr, a.ctypes.data, a.size * 4, None))
        self.ctx.check(self.ctx._lib.sfn_synchronize(self.ctx.ptr, None))

    def download(self, n=None) -> np.ndarray:
        n = self.n if n is None else n
        out = np.empty(n, dtype=np.uint32)
        self.ctx.check(self.ctx._lib.sfn_download(self.ctx.ptr, out.ctypes.data, self.ptr, n * 4, None))
        self.ctx.check(self.ctx._lib.sfn_synchronize(self.ctx.ptr, None))
        return out

    def free(self):
        if self.ptr:
            self.ctx._lib.sfn_buf_free(self.ctx.ptr, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------- backends
def _log2_exact(n: int) -> int:
    if n == 0 or n & (n - 1):
        raise InvalidSize(n)
    return n.bit_length() - 1


class _GpuBackend:
    FIELD = L.FIELD_M31
    VARIANT = L.VARIANT_AUTO
    NAME = ""
    MODULUS = 0x7FFFFFFF

    def __init__(self, shader_dir=None, ctx: HipContext | None = None):
        # shader_dir is accepted for signature compatibility; kernels are built in.
        self.ctx = ctx or default_context()

    def name(self) -> str:
        return self.NAME

    # src/ntt/metal_r2.rs:69-114 / :117-173
    def _gpu(self, direction, data, log_n):
        a = _u32(data)
        if a.size != (1 << log_n) or log_n > 30:
            raise InvalidSize(a.size)
        if log_n == 0:
            return a.copy(), 0
        return self.ctx.ntt(self.FIELD, self.VARIANT, direction, log_n, 1, a)

    def forward_ntt_gpu(self, data, log_n: int):
        return self._gpu(L.FORWARD, data, log_n)

    def inverse_ntt_gpu(self, data, log_n: int):
        return self._gpu(L.INVERSE, data, log_n)

    # NttBackend trait: in place on a u32 numpy array (the raw field values)
    def forward_ntt(self, data: np.ndarray, twiddles=None) -> np.ndarray:
        log_n = _log2_exact(len(data))
        if log_n == 0:
            return data
        out, _ = self.forward_ntt_gpu(data, log_n)
        data[...] = out
        return data

    def inverse_ntt(self, data: np.ndarray, twiddles=None) -> np.ndarray:
        log_n = _log2_exact(len(data))
        if log_n == 0:
            return data
        out, _ = self.inverse_ntt_gpu(data, log_n)
        data[...] = out
        return data

    def pointwise_mul(self, a, b, out: np.ndarray) -> np.ndarray:
        if len(a) != len(b) or len(a) != len(out):
            raise InvalidSize(len(a))
        out[...] = self.ctx.pointwise_mul(self.FIELD, a, b)
        return out


# ---- M31 circle NTT variants (src/ntt/m31/*.rs, src/ntt/metal_r2.rs:337) ----
class CtDitR2(_GpuBackend):
    NAME, VARIANT = "hip-ct-dit-r2", L.CT_DIT_R2


class CtGsR2(_GpuBackend):
    NAME, VARIANT = "hip-ct-gs-r2", L.CT_GS_R2


class StockhamR2(_GpuBackend):
    NAME, VARIANT = "hip-stockham-r2", L.STOCKHAM_R2


class CtGsR4(_GpuBackend):
    NAME, VARIANT = "hip-ct-gs-r4", L.CT_GS_R4


class TunedM31(_GpuBackend):
    NAME, VARIANT = "hip-tuned", L.VARIANT_AUTO


# ---- BabyBear variants (src/ntt/babybear/*.rs) ----
class _Bb(_GpuBackend):
    FIELD = L.FIELD_BABYBEAR
    MODULUS = 0x78000001
    MAX_LOG = 27  # two-adicity (src/field/babybear.rs:70)

    def _gpu(self, direction, data, log_n):
        a = _u32(data)
        if a.size != (1 << log_n) or log_n > self.MAX_LOG:
            raise InvalidSize(a.size)
        if log_n == 0:
            return a.copy(), 0
        return self.ctx.ntt(self.FIELD, self.VARIANT, direction, log_n, 1, a)


class BbCtDitR2(_Bb):
    NAME, VARIANT = "bb-hip-ct-dit-r2", L.CT_DIT_R2


class BbR2(_Bb):
    NAME, VARIANT = "bb-hip-r2", L.CT_GS_R2


class BbStockhamR2(_Bb):
    NAME, VARIANT = "bb-hip-stockham-r2", L.STOCKHAM_R2


class BbCtGsR4(_Bb):
    NAME, VARIANT = "bb-hip-ct-gs-r4", L.CT_GS_R4


class BbBatchNtt(_Bb):
    """Batched BabyBear NTT, SoA layout data[col * n + row] (src/ntt/babybear/batch.rs)."""
    NAME, VARIANT = "bb-hip-batch", L.VARIANT_AUTO

    def _batch(self, direction, data, log_n, batch_size, layout):
        a = _u32(data)
        if a.size != (batch_size << log_n) or log_n > self.MAX_LOG or batch_size == 0:
            raise InvalidSize(a.size)
        if log_n == 0:
            return a.copy(), 0
        return self.ctx.ntt(self.FIELD, self.VARIANT, direction, log_n, batch_size, a, layout)

    def forward_ntt_batch_gpu(self, data, log_n: int, batch_size: int, layout=L.LAYOUT_SOA):
        return self._batch(L.FORWARD, data, log_n, batch_size, layout)

    def inverse_ntt_batch_gpu(self, data, log_n: int, batch_size: int, layout=L.LAYOUT_SOA):
        return self._batch(L.INVERSE, data, log_n, batch_size, layout)


# ---- KoalaBear variants (README.md:20 lists the field as planned) ----
# Same conventions as BabyBear: Montgomery raw values (R = 2^32), forward from
# natural order to bit-reversed evaluations, inverse scaled by 1/n; p = 2^31 -
# 2^24 + 1, two-adicity 24.
class _Kb(_Bb):
    FIELD = L.FIELD_KOALABEAR
    MODULUS = 0x7F000001
    MAX_LOG = 24


class KbCtDitR2(_Kb):
    NAME, VARIANT = "kb-hip-ct-dit-r2", L.CT_DIT_R2


class KbR2(_Kb):
    NAME, VARIANT = "kb-hip-r2", L.CT_GS_R2


class KbStockhamR2(_Kb):
    NAME, VARIANT = "kb-hip-stockham-r2", L.STOCKHAM_R2


class KbCtGsR4(_Kb):
    NAME, VARIANT = "kb-hip-ct-gs-r4", L.CT_GS_R4


class KbBatchNtt(_Kb, BbBatchNtt):
    """Batched KoalaBear NTT, SoA layout data[col * n + row] (BbBatchNtt's API)."""
    NAME, VARIANT = "kb-hip-batch", L.VARIANT_AUTO


class LdeResult:
    """src/lde/mod.rs:19-25: data in SoA layout data[col * n_ext + row], device time."""

    def __init__(self, data: np.ndarray, total_ns: int):
        self.data, self.total_ns = data, total_ns


class CosetLdeBatch:
    """Batched BabyBear coset LDE (src/lde/mod.rs:15-113)."""

    def __init__(self, shader_dir=None, ctx: HipContext | None = None):
        self.ctx = ctx or default_context()

    def execute(self, data, log_n: int, batch_size: int, added_bits: int, shift_raw: int = 0,
                layout: int = L.LAYOUT_SOA) -> LdeResult:
        a = _u32(data)
        if a.size != (batch_size << log_n):  # src/lde/mod.rs:54-56
            raise InvalidSize(a.size)
        out, ns = self.ctx.coset_lde(a, log_n, batch_size, added_bits, shift_raw, layout)
        return LdeResult(out, ns)

    def execute_natural(self, data, log_n: int, batch_size: int, added_bits: int, shift_raw: int = 0,
                        layout: int = L.LAYOUT_SOA) -> LdeResult:
        """The same LDE with natural row order on both sides (input row i = P(g_n^i),
        output row k = P(shift * g_N^k)): the order the reference converts to with
        bit_reverse_permutation to compare against Plonky3 (tests/ntt_cross_validation.rs:21-36)."""
        a = _u32(data)
        if a.size != (batch_size << log_n):
            raise InvalidSize(a.size)
        out, ns = self.ctx.coset_lde(a, log_n, batch_size, added_bits, shift_raw, layout, natural=True)
        return LdeResult(out, ns)

    def device_info(self) -> dict:
        return self.ctx.device_info()


# reference type names
MetalCtDitR2, MetalCtGsR2, MetalStockhamR2, MetalCtGsR4 = CtDitR2, CtGsR2, StockhamR2, CtGsR4
BbMetalCtDitR2, BbMetalR2, BbMetalStockhamR2, BbMetalCtGsR4 = BbCtDitR2, BbR2, BbStockhamR2, BbCtGsR4
KbMetalCtDitR2, KbMetalR2, KbMetalStockhamR2, KbMetalCtGsR4 = KbCtDitR2, KbR2, KbStockhamR2, KbCtGsR4

M31_VARIANTS = [CtDitR2, CtGsR2, StockhamR2, CtGsR4, TunedM31]
BB_VARIANTS = [BbCtDitR2, BbR2, BbStockhamR2, BbCtGsR4, BbBatchNtt]
KB_VARIANTS = [KbCtDitR2, KbR2, KbStockhamR2, KbCtGsR4, KbBatchNtt]
