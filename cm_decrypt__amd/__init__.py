"""MI355X Classic McEliece batched decryption engine (Python binding of the C ABI).

The compute path is the HIP library ``libcmdecrypt.so`` built next to this file
(``cm_decrypt__amd/Makefile``); this module only loads it with ctypes and moves
buffers.  Importing fails loudly when the library is missing: there is no Python
or CPU fallback for decryption.

Reference interface mirrored here (paths relative to the reference repository):
  * ``initialisation`` / ``compute_inverses`` -> :meth:`Context.load_key`
    (CPU/param/param6960119/src/util.c:60-100, GPU InitializeC in
    GPU_Baseline/param/param8192128/include/decrypt.h:147-173);
  * ``synd`` / ``bm`` / ``root`` -> :meth:`Context.synd`, :meth:`Context.bm`,
    :meth:`Context.root` (util.c:102-114, CPU/src/common/bm.c:12-68,
    CPU/src/common/root_13.c:233-240);
  * the batch drivers ``decrypt``/``decrypt_all`` -> :meth:`Context.decrypt`
    (GPU_Baseline/param/param8192128/src/cuda/Decrypt.cu:177-324);
  * ``write_error_positions`` -> :func:`errorstream` (decrypt.c:55-66).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CMD_DECRYPT_LIB selects another build of the same library (the sanitizer build of the host
# tests); it must export the same C ABI.
LIB_PATH = os.environ.get("CMD_DECRYPT_LIB") or os.path.join(_HERE, "libcmdecrypt.so")

PARAM_IDS = (348864, 460896, 6688128, 6960119, 8192128)
MODE_PLAIN = 0
MODE_FUSED = 1
MODE_KEM = 2           # rows out are 32-byte session keys (crypto_kem_dec)
KEM_KEY_BYTES = 32
MODE_POSITIONS = 3     # rows out are t uint16 error positions, ascending, 0xFFFF-padded
NOALLOC_MAX_ITEMS = 262144  # items per cmd_decrypt_device_noalloc call

CMD_OK = 0
CMD_ERR_FILE = -1
CMD_ERR_CRYPTO = -4
CMD_ERR_PARAM = -10
CMD_ERR_ARG = -11
CMD_ERR_HIP = -12
CMD_ERR_NODEV = -13
CMD_ERR_BUDGET = -14
CMD_ERR_KEY = -15
CMD_ERR_FULL = -16
CMD_ERR_NOMEM = -17
CMD_ERR_SINGULAR = -18
CMD_ERR_PENDING = -19
CMD_ERR_TIMEOUT = -20


class CmdError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        msg = _lib.cmd_strerror(code).decode()
        detail = _lib.cmd_last_error().decode()
        super().__init__(f"{where}: {msg} ({code}){': ' + detail if detail else ''}")


def _load():
    # torch (when installed) bundles its own HIP runtime under the same soname; loading it
    # first makes this library bind to that one runtime instead of a second copy.
    if not os.environ.get("CMD_DECRYPT_NO_TORCH"):
        try:
            import torch  # noqa: F401
        except Exception:
            pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C {_HERE}` (or __graft_entry__.build())")
    return ctypes.CDLL(LIB_PATH)


_lib = _load()

_c = ctypes
_P = _c.c_void_p
_SZ = _c.c_size_t


class ParamInfo(_c.Structure):
    _fields_ = [(n, _c.c_int32) for n in
                ("id", "m", "n", "t", "ct_bytes", "sb", "sk_bytes", "cond_bytes", "e_bytes")]


class Schedule(_c.Structure):
    _fields_ = [(n, _c.c_int32) for n in ("tile_cts", "waves_per_group", "waves_per_ct", "groups_per_cu",
                                          "root_alg")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class ServerConfig(_c.Structure):
    _fields_ = [("max_batch", _c.c_int32), ("max_inflight", _c.c_int32), ("queue_capacity", _c.c_int32),
                ("min_batch", _c.c_int32), ("max_wait_us", _c.c_double), ("io_budget_bytes", _c.c_uint64)]


class ServerStats(_c.Structure):
    _fields_ = [("submitted", _c.c_uint64), ("completed", _c.c_uint64), ("rejected", _c.c_uint64),
                ("launches", _c.c_uint64), ("failed", _c.c_uint64), ("mean_batch", _c.c_double),
                ("p50_us", _c.c_double), ("p99_us", _c.c_double), ("p999_us", _c.c_double),
                ("max_us", _c.c_double), ("queue_p99_us", _c.c_double), ("queue_max_us", _c.c_double),
                ("device_p99_us", _c.c_double), ("device_max_us", _c.c_double), ("gpu_p99_us", _c.c_double),
                ("gpu_max_us", _c.c_double), ("busy_s", _c.c_double),
                ("step_launch_max_us", _c.c_double), ("step_retire_max_us", _c.c_double),
                ("step_poll_max_us", _c.c_double), ("wait_overshoot_max_us", _c.c_double)]

    def as_dict(self):
        return {n: (float(getattr(self, n)) if t is _c.c_double else int(getattr(self, n)))
                for n, t in self._fields_}


def _sig(name, restype, *args):
    f = getattr(_lib, name)
    f.restype = restype
    f.argtypes = list(args)
    return f


_I = _c.c_int
_I32 = _c.c_int32
_U64 = _c.c_uint64
_sig("cmd_strerror", _c.c_char_p, _I)
_sig("cmd_last_error", _c.c_char_p)
_sig("cmd_param_get", _I, _I32, _c.POINTER(ParamInfo))
_sig("cmd_support_gen", _I, _I32, _P, _P)
_sig("cmd_key_expand_host", _I, _I32, _P, _SZ, _P, _P, _P)
_sig("cmd_errorstream_format", _I, _I32, _P, _SZ, _SZ, _P, _SZ, _c.POINTER(_SZ))
_sig("cmd_errorstream_write", _I, _I32, _c.c_char_p, _P, _SZ, _SZ, _I)
_sig("cmd_gen_keypair", _I, _I32, _U64, _P)
_sig("cmd_gen_encrypt", _I, _I32, _P, _SZ, _U64, _SZ, _I32, _P, _P)
_sig("cmd_gen_keys", _I, _I32, _U64, _SZ, _P, _I)
_sig("cmd_gen_keys_device", _I, _P, _I32, _U64, _SZ, _P, _I)
_sig("cmd_controlbits", _I, _I32, _P, _P)
_sig("cmd_ctx_create", _I, _I32, _U64, _c.POINTER(_P))
_sig("cmd_ctx_destroy", _I, _P)
_sig("cmd_ctx_memory", _I, _P, _c.POINTER(_U64), _c.POINTER(_U64))
_sig("cmd_ctx_set_budget", _I, _P, _U64)
_sig("cmd_key_load", _I, _P, _I32, _P, _SZ, _c.POINTER(_I32))
_sig("cmd_key_unload", _I, _P, _I32)
_sig("cmd_key_device_bytes", _I, _I32, _c.POINTER(_U64))
_sig("cmd_key_export", _I, _P, _I32, _P, _P, _P, _P, _c.POINTER(_I32))
_sig("cmd_decrypt", _I, _P, _I32, _SZ, _P, _P, _SZ, _P, _SZ)
_sig("cmd_decrypt_device", _I, _P, _I32, _SZ, _P, _P, _SZ, _P, _SZ, _P)
_sig("cmd_decrypt_device_noalloc", _I, _P, _SZ, _P, _P, _SZ, _P, _SZ, _P)
_sig("cmd_batch_create", _I, _P, _I32, _SZ, _P, _SZ, _SZ, _I32, _c.POINTER(_P))
_sig("cmd_batch_run", _I, _P, _P, _P, _P)
_sig("cmd_batch_destroy", _I, _P)
_sig("cmd_stage_synd", _I, _P, _I32, _SZ, _P, _SZ, _P)
_sig("cmd_stage_synd_full", _I, _P, _I32, _SZ, _P, _SZ, _P)
_sig("cmd_kem_dec", _I, _P, _SZ, _P, _P, _SZ, _P, _SZ)
_sig("cmd_host_alloc", _I, _P, _SZ, _c.POINTER(_P))
_sig("cmd_gen_encrypt_device", _I, _P, _SZ, _P, _P, _SZ, _I32, _P, _P)
_sig("cmd_errorstream_format_positions", _I, _I32, _P, _SZ, _SZ, _P, _SZ, _c.POINTER(_SZ))
_sig("cmd_host_free", _I, _P, _P)
_sig("cmd_stage_bm", _I, _P, _I32, _SZ, _P, _P)
_sig("cmd_stage_root", _I, _P, _I32, _SZ, _P, _P, _SZ)
_sig("cmd_schedule_get", _I, _P, _I32, _c.POINTER(Schedule))
_sig("cmd_schedule_set", _I, _P, _I32, _c.POINTER(Schedule))
_sig("cmd_set_tile_fit", _I, _P, _I32)
_sig("cmd_autotune", _I, _P, _I32, _SZ, _P, _I32, _c.POINTER(Schedule), _c.POINTER(_c.c_double))
_sig("cmd_schedule_save", _I, _P, _c.c_char_p)
_sig("cmd_schedule_load", _I, _P, _c.c_char_p)
_sig("cmd_server_create", _I, _P, _c.POINTER(ServerConfig), _c.POINTER(_P))
_sig("cmd_server_submit", _I, _P, _I32, _P, _c.POINTER(_U64))
_sig("cmd_server_submit_batch", _I, _P, _SZ, _P, _P, _SZ, _P, _c.POINTER(_SZ))
_sig("cmd_server_submit_batch_mode", _I, _P, _I32, _SZ, _P, _P, _SZ, _P, _c.POINTER(_SZ))
_sig("cmd_server_result_bytes", _I, _P, _U64, _c.POINTER(_SZ))
_sig("cmd_server_wait_batch", _I, _P, _SZ, _P, _P, _SZ, _c.c_double)
_sig("cmd_server_poll", _I, _P, _U64, _P, _SZ)
_sig("cmd_server_wait", _I, _P, _U64, _P, _SZ, _c.c_double)
_sig("cmd_server_latencies", _I, _P, _P, _P, _SZ, _c.POINTER(_SZ))
_sig("cmd_server_stats_get", _I, _P, _c.POINTER(ServerStats))
_sig("cmd_server_launch_log", _I, _P, _P, _SZ, _c.POINTER(_SZ))
_sig("cmd_server_reset_stats", _I, _P)
_sig("cmd_server_destroy", _I, _P)
_sig("cmd_set_kernel_timing", _I, _P, _I32)
_sig("cmd_kernel_timing", _I, _P, _c.POINTER(_c.c_double), _c.POINTER(_U64))
_sig("cmd_valu_peak", _I, _P, _c.POINTER(_c.c_double))
_sig("cmd_valu_issue_rate", _I, _P, _I32, _I32, _c.POINTER(_c.c_double))
_sig("cmd_chain_latency", _I, _P, _c.POINTER(_c.c_double))
_sig("cmd_build_info", _c.c_char_p)

lib = _lib


def _check(rc: int, where: str):
    if rc != CMD_OK:
        raise CmdError(rc, where)


def _check_rows(buf: np.ndarray, count: int, stride: int, name: str):
    """The library reads `count` rows `stride` bytes apart: refuse a buffer that is smaller."""
    if count and buf.nbytes < count * stride:
        raise ValueError(f"{name}: {buf.nbytes} bytes for {count} rows of {stride}")


def _ptr(a) -> Optional[int]:
    if a is None:
        return None
    return a.ctypes.data


@dataclass(frozen=True)
class Params:
    id: int
    m: int
    n: int
    t: int
    ct_bytes: int
    sb: int
    sk_bytes: int
    cond_bytes: int
    e_bytes: int


def params(param_id: int) -> Params:
    info = ParamInfo()
    _check(_lib.cmd_param_get(param_id, _c.byref(info)), "cmd_param_get")
    return Params(*(int(getattr(info, f)) for f, _ in ParamInfo._fields_))


def build_info() -> str:
    return _lib.cmd_build_info().decode()


# ---- host-only helpers -------------------------------------------------------------------

def key_device_bytes(param_id: int) -> int:
    """Device bytes one resident key of this parameter set occupies."""
    b = _U64(0)
    _check(_lib.cmd_key_device_bytes(param_id, _c.byref(b)), "cmd_key_device_bytes")
    return int(b.value)


def support_gen(param_id: int, cond: np.ndarray) -> np.ndarray:
    p = params(param_id)
    cond = np.ascontiguousarray(cond, dtype=np.uint8)
    if cond.size < p.cond_bytes:
        raise ValueError("condition bits too short")
    L = np.zeros(p.n, np.uint16)
    _check(_lib.cmd_support_gen(param_id, _ptr(cond), _ptr(L)), "cmd_support_gen")
    return L


def key_expand_host(param_id: int, sk: np.ndarray):
    """(g[t+1], L[n], table[sb, 2t]) exactly as initialisation() + compute_inverses()."""
    p = params(param_id)
    sk = np.ascontiguousarray(sk, dtype=np.uint8)
    g = np.zeros(p.t + 1, np.uint16)
    L = np.zeros(p.n, np.uint16)
    table = np.zeros((p.sb, 2 * p.t), np.uint16)
    _check(_lib.cmd_key_expand_host(param_id, _ptr(sk), sk.size, _ptr(g), _ptr(L), _ptr(table)),
           "cmd_key_expand_host")
    return g, L, table


def errorstream(param_id: int, e: np.ndarray) -> bytes:
    """Text of the reference errorstream for error vectors e (count x n/8 bytes)."""
    p = params(param_id)
    e = np.ascontiguousarray(e, dtype=np.uint8).reshape(-1, p.e_bytes)
    need = _SZ(0)
    _check(_lib.cmd_errorstream_format(param_id, _ptr(e), p.e_bytes, e.shape[0], None, 0, _c.byref(need)),
           "cmd_errorstream_format")
    buf = _c.create_string_buffer(need.value)
    _check(_lib.cmd_errorstream_format(param_id, _ptr(e), p.e_bytes, e.shape[0], _c.addressof(buf), need.value,
                                       _c.byref(need)), "cmd_errorstream_format")
    return buf.raw[:need.value]


def errorstream_positions(param_id: int, pos: np.ndarray) -> bytes:
    """The same text from position rows (count x t uint16, 0xFFFF-padded; CMD_MODE_POSITIONS)."""
    p = params(param_id)
    pos = np.ascontiguousarray(pos, dtype=np.uint16).reshape(-1, p.t)
    need = _SZ(0)
    _check(_lib.cmd_errorstream_format_positions(param_id, _ptr(pos), 2 * p.t, pos.shape[0], None, 0,
                                                 _c.byref(need)), "cmd_errorstream_format_positions")
    buf = _c.create_string_buffer(max(1, need.value))
    _check(_lib.cmd_errorstream_format_positions(param_id, _ptr(pos), 2 * p.t, pos.shape[0], _c.addressof(buf),
                                                 need.value, _c.byref(need)), "cmd_errorstream_format_positions")
    return buf.raw[:need.value]


def errorstream_write(param_id: int, path: str, e: np.ndarray, append: bool = False):
    p = params(param_id)
    e = np.ascontiguousarray(e, dtype=np.uint8).reshape(-1, p.e_bytes)
    _check(_lib.cmd_errorstream_write(param_id, path.encode(), _ptr(e), p.e_bytes, e.shape[0], int(append)),
           "cmd_errorstream_write")


# ---- generator -----------------------------------------------------------------------------

def gen_keypair(param_id: int, seed: int) -> np.ndarray:
    p = params(param_id)
    sk = np.zeros(p.sk_bytes, np.uint8)
    _check(_lib.cmd_gen_keypair(param_id, seed, _ptr(sk)), "cmd_gen_keypair")
    return sk


def gen_keys(param_id: int, base_seed: int, n_keys: int, threads: int = 8) -> np.ndarray:
    p = params(param_id)
    out = np.zeros((n_keys, p.sk_bytes), np.uint8)
    _check(_lib.cmd_gen_keys(param_id, base_seed, n_keys, _ptr(out), threads), "cmd_gen_keys")
    return out


def controlbits(m: int, pi: np.ndarray) -> np.ndarray:
    """Benes condition bits whose network maps the identity to permutation pi of 2^m."""
    pi = np.ascontiguousarray(pi, dtype=np.int16)
    if pi.size != (1 << m):
        raise ValueError("pi must have 2^m entries")
    cond = np.zeros((2 * m - 1) * (1 << (m - 4)), np.uint8)
    _check(_lib.cmd_controlbits(m, _ptr(pi), _ptr(cond)), "cmd_controlbits")
    return cond


def gen_encrypt(param_id: int, sk: np.ndarray, seed: int, count: int, weight: int = -1):
    """(ct[count, ct_bytes], e[count, n/8]) with errors of exactly `weight` (default t)."""
    p = params(param_id)
    sk = np.ascontiguousarray(sk, dtype=np.uint8)
    ct = np.zeros((count, p.ct_bytes), np.uint8)
    e = np.zeros((count, p.e_bytes), np.uint8)
    _check(_lib.cmd_gen_encrypt(param_id, _ptr(sk), sk.size, seed, count, weight, _ptr(ct), _ptr(e)),
           "cmd_gen_encrypt")
    return ct, e


# ---- GPU context -----------------------------------------------------------------------------

class Context:
    """One GPU, its resident keys and work buffers (cmd_ctx)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        h = _P()
        _check(_lib.cmd_ctx_create(device, hbm_budget_bytes, _c.byref(h)), "cmd_ctx_create")
        self._h = h
        self.key_param = {}

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            _lib.cmd_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def memory(self):
        used, budget = _U64(0), _U64(0)
        _check(_lib.cmd_ctx_memory(self._h, _c.byref(used), _c.byref(budget)), "cmd_ctx_memory")
        return int(used.value), int(budget.value)

    def set_budget(self, hbm_budget_bytes: int):
        """New HBM budget for later allocations (keys, work buffers, server scratch)."""
        _check(_lib.cmd_ctx_set_budget(self._h, hbm_budget_bytes), "cmd_ctx_set_budget")

    def load_key(self, param_id: int, sk: np.ndarray) -> int:
        sk = np.ascontiguousarray(sk, dtype=np.uint8)
        kid = _I32(-1)
        _check(_lib.cmd_key_load(self._h, param_id, _ptr(sk), sk.size, _c.byref(kid)), "cmd_key_load")
        self.key_param[int(kid.value)] = param_id
        return int(kid.value)

    def export_key(self, key_id: int):
        """Device arrays of a resident key: (table sb x 2t, L n, lbits m x stride, benes (2m-1) x 2^(m-5))."""
        p = params(self.key_param[key_id])
        stride = _I32(0)
        _check(_lib.cmd_key_export(self._h, key_id, None, None, None, None, _c.byref(stride)), "cmd_key_export")
        table = np.zeros((p.sb, 2 * p.t), np.uint16)
        L = np.zeros(p.n, np.uint16)
        lbits = np.zeros((p.m, stride.value), np.uint32)
        benes = np.zeros((2 * p.m - 1, 1 << (p.m - 5)), np.uint32)
        _check(_lib.cmd_key_export(self._h, key_id, _ptr(table), _ptr(L), _ptr(lbits), _ptr(benes), None),
               "cmd_key_export")
        return table, L, lbits, benes

    def unload_key(self, key_id: int):
        _check(_lib.cmd_key_unload(self._h, key_id), "cmd_key_unload")
        self.key_param.pop(key_id, None)

    def decrypt(self, key_ids: Sequence[int], ct: np.ndarray, mode: int = MODE_FUSED,
                ct_stride: Optional[int] = None, e_stride: Optional[int] = None,
                out: Optional[np.ndarray] = None) -> np.ndarray:
        """Error vectors for ciphertext rows `ct` (count x ct_stride), item i under key_ids[i].

        Returns count x e_stride bytes (e_stride defaults to the largest n/8 of the keys), in
        `out` when given (a C-contiguous uint8 array, e.g. from host_array)."""
        keys = np.ascontiguousarray(key_ids, dtype=np.int32)
        ct = np.ascontiguousarray(ct, dtype=np.uint8)
        count = keys.size
        if ct_stride is None:
            ct_stride = ct.shape[1] if ct.ndim == 2 else ct.size // max(count, 1)
        if e_stride is None:
            # keys this object did not load (or unloaded) get the widest row; the library
            # itself rejects unknown keys with CMD_ERR_KEY
            e_stride = max((params(self.key_param[int(k)]).e_bytes if int(k) in self.key_param else 1024)
                           for k in np.unique(keys)) if count else 0
        _check_rows(ct, count, ct_stride, "ct")
        if out is None:
            e = np.zeros((count, e_stride), np.uint8)
        else:
            if out.dtype != np.uint8 or not out.flags.c_contiguous or out.size < count * e_stride:
                raise ValueError("out must be a C-contiguous uint8 array of count x e_stride bytes")
            e = out
        _check(_lib.cmd_decrypt(self._h, mode, count, _ptr(keys), _ptr(ct), ct_stride, _ptr(e), e_stride),
               "cmd_decrypt")
        return e

    def positions(self, key_ids: Sequence[int], ct: np.ndarray, ct_stride: Optional[int] = None) -> np.ndarray:
        """Error positions (count x t_max uint16, ascending, 0xFFFF-padded) of CMD_MODE_POSITIONS."""
        keys = np.ascontiguousarray(key_ids, dtype=np.int32)
        t = max((params(self.key_param[int(k)]).t if int(k) in self.key_param else 128)
                for k in np.unique(keys)) if keys.size else 1
        rows = self.decrypt(keys, ct, mode=MODE_POSITIONS, ct_stride=ct_stride, e_stride=2 * t)
        return rows.view(np.uint16)

    def gen_encrypt(self, key_ids: Sequence[int], seeds: Sequence[int], per_key: int, weight: int = -1):
        """gen_encrypt(param, sk_k, seeds[k], per_key, weight) for every resident key k (one set),
        computed on the GPU: (ct, e), key-major rows."""
        keys = np.ascontiguousarray(key_ids, dtype=np.int32)
        sd = np.ascontiguousarray(seeds, dtype=np.uint64)
        p = params(self.key_param[int(keys[0])])
        ct = np.zeros((keys.size * per_key, p.ct_bytes), np.uint8)
        e = np.zeros((keys.size * per_key, p.e_bytes), np.uint8)
        _check(_lib.cmd_gen_encrypt_device(self._h, keys.size, _ptr(keys), _ptr(sd), per_key, weight, _ptr(ct),
                                           _ptr(e)), "cmd_gen_encrypt_device")
        return ct, e

    def gen_keys(self, param_id: int, base_seed: int, n_keys: int, threads: int = 8) -> np.ndarray:
        """The key pairs of gen_keys (same bytes), with each attempt's systematic-form test on the GPU."""
        p = params(param_id)
        out = np.zeros((n_keys, p.sk_bytes), np.uint8)
        _check(_lib.cmd_gen_keys_device(self._h, param_id, base_seed, n_keys, _ptr(out), threads),
               "cmd_gen_keys_device")
        return out

    def host_array(self, shape, dtype=np.uint8) -> np.ndarray:
        """A numpy array in pinned host memory (cmd_host_alloc): decrypt() moves such buffers by
        DMA without staging.  Valid until host_free() or until the context closes."""
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = _P()
        _check(_lib.cmd_host_alloc(self._h, nbytes, _c.byref(p)), "cmd_host_alloc")
        buf = (_c.c_uint8 * nbytes).from_address(p.value)
        return np.frombuffer(buf, dtype=dtype).reshape(shape)

    def host_free(self, arr: np.ndarray):
        _check(_lib.cmd_host_free(self._h, arr.ctypes.data), "cmd_host_free")

    def kem_dec(self, key_ids: Sequence[int], ct: np.ndarray, ct_stride: Optional[int] = None) -> np.ndarray:
        """Session keys (count x 32 bytes) of crypto_kem_dec for ciphertext rows `ct`."""
        keys = np.ascontiguousarray(key_ids, dtype=np.int32)
        ct = np.ascontiguousarray(ct, dtype=np.uint8)
        count = keys.size
        if ct_stride is None:
            ct_stride = ct.shape[1] if ct.ndim == 2 else ct.size // max(count, 1)
        _check_rows(ct, count, ct_stride, "ct")
        out = np.zeros((count, KEM_KEY_BYTES), np.uint8)
        _check(_lib.cmd_kem_dec(self._h, count, _ptr(keys), _ptr(ct), ct_stride, _ptr(out), KEM_KEY_BYTES),
               "cmd_kem_dec")
        return out

    def decrypt_device(self, key_ids, d_ct: int, ct_stride: int, d_e: int, e_stride: int,
                       mode: int = MODE_FUSED, stream: Optional[int] = None):
        keys = np.ascontiguousarray(key_ids, dtype=np.int32)
        _check(_lib.cmd_decrypt_device(self._h, mode, keys.size, _ptr(keys), d_ct, ct_stride, d_e, e_stride,
                                       stream), "cmd_decrypt_device")

    # reference stage functions
    def synd(self, key_id: int, ct: np.ndarray) -> np.ndarray:
        p = params(self.key_param[key_id])
        ct = np.ascontiguousarray(ct, dtype=np.uint8).reshape(-1, p.ct_bytes)
        s = np.zeros((ct.shape[0], 2 * p.t), np.uint16)
        _check(_lib.cmd_stage_synd(self._h, key_id, ct.shape[0], _ptr(ct), p.ct_bytes, _ptr(s)), "cmd_stage_synd")
        return s

    def synd_full(self, key_id: int, r: np.ndarray) -> np.ndarray:
        """s_j(r) = sum_{i<n} r_i L_i^j / g(L_i)^2, j < 2t, for rows r of n/8 bytes."""
        p = params(self.key_param[key_id])
        r = np.ascontiguousarray(r, dtype=np.uint8).reshape(-1, p.e_bytes)
        s = np.zeros((r.shape[0], 2 * p.t), np.uint16)
        _check(_lib.cmd_stage_synd_full(self._h, key_id, r.shape[0], _ptr(r), p.e_bytes, _ptr(s)),
               "cmd_stage_synd_full")
        return s

    def bm(self, param_id: int, s: np.ndarray) -> np.ndarray:
        p = params(param_id)
        s = np.ascontiguousarray(s, dtype=np.uint16).reshape(-1, 2 * p.t)
        loc = np.zeros((s.shape[0], p.t + 1), np.uint16)
        _check(_lib.cmd_stage_bm(self._h, param_id, s.shape[0], _ptr(s), _ptr(loc)), "cmd_stage_bm")
        return loc

    def root(self, key_id: int, loc: np.ndarray) -> np.ndarray:
        p = params(self.key_param[key_id])
        loc = np.ascontiguousarray(loc, dtype=np.uint16).reshape(-1, p.t + 1)
        e = np.zeros((loc.shape[0], p.e_bytes), np.uint8)
        _check(_lib.cmd_stage_root(self._h, key_id, loc.shape[0], _ptr(loc), _ptr(e), p.e_bytes),
               "cmd_stage_root")
        return e

    # schedules
    def schedule(self, param_id: int) -> dict:
        s = Schedule()
        _check(_lib.cmd_schedule_get(self._h, param_id, _c.byref(s)), "cmd_schedule_get")
        return s.as_dict()

    def set_schedule(self, param_id: int, **kw):
        s = Schedule(**self.schedule(param_id))
        for k, v in kw.items():
            setattr(s, k, v)
        _check(_lib.cmd_schedule_set(self._h, param_id, _c.byref(s)), "cmd_schedule_set")

    def set_tile_fit(self, on: bool):
        """Narrow tiles of batches too small to fill the GPU (default on)."""
        _check(_lib.cmd_set_tile_fit(self._h, int(on)), "cmd_set_tile_fit")

    def autotune(self, param_id: int, key_ids, reps: int = 3):
        keys = np.ascontiguousarray(key_ids, dtype=np.int32)
        best = Schedule()
        us = _c.c_double(0)
        _check(_lib.cmd_autotune(self._h, param_id, keys.size, _ptr(keys), reps, _c.byref(best), _c.byref(us)),
               "cmd_autotune")
        return best.as_dict(), float(us.value)

    def save_schedules(self, path: str):
        _check(_lib.cmd_schedule_save(self._h, path.encode()), "cmd_schedule_save")

    def load_schedules(self, path: str):
        _check(_lib.cmd_schedule_load(self._h, path.encode()), "cmd_schedule_load")

    def set_kernel_timing(self, on: bool):
        _check(_lib.cmd_set_kernel_timing(self._h, int(on)), "cmd_set_kernel_timing")

    def valu_peak(self) -> float:
        """Measured VALU issue rate (32-bit lane operations per second)."""
        v = _c.c_double(0)
        _check(_lib.cmd_valu_peak(self._h, _c.byref(v)), "cmd_valu_peak")
        return float(v.value)

    def valu_issue_rate(self, waves_per_simd: int, chains: int) -> float:
        """VALU wave instructions per SIMD per clock with waves_per_simd wavefronts per SIMD
        running `chains` independent v_xor_b32 chains per lane."""
        v = _c.c_double(0)
        _check(_lib.cmd_valu_issue_rate(self._h, waves_per_simd, chains, _c.byref(v)), "cmd_valu_issue_rate")
        return float(v.value)

    def chain_latency(self):
        """Dependent-step latencies of one wavefront in shader-clock ticks: (v_xor_b32, DPP row
        shift + v_xor_b32, v_readlane + scalar + v_xor_b32)."""
        t = (_c.c_double * 3)()
        _check(_lib.cmd_chain_latency(self._h, t), "cmd_chain_latency")
        return tuple(float(x) for x in t)

    def kernel_timing(self):
        us, n = _c.c_double(0), _U64(0)
        _check(_lib.cmd_kernel_timing(self._h, _c.byref(us), _c.byref(n)), "cmd_kernel_timing")
        return float(us.value), int(n.value)


class Batch:
    """A planned batch (fixed key assignment) replayed with cmd_batch_run, optionally as a hipGraph."""

    def __init__(self, ctx: Context, key_ids, ct_stride: int, e_stride: int, mode: int = MODE_FUSED,
                 use_graph: bool = True):
        keys = np.ascontiguousarray(key_ids, dtype=np.int32)
        h = _P()
        _check(_lib.cmd_batch_create(ctx.handle, mode, keys.size, _ptr(keys), ct_stride, e_stride, int(use_graph),
                                     _c.byref(h)), "cmd_batch_create")
        self._h = h
        self.ctx = ctx
        self.count = keys.size

    def run(self, d_ct: int, d_e: int, stream: Optional[int] = None):
        _check(_lib.cmd_batch_run(self._h, d_ct, d_e, stream), "cmd_batch_run")

    def close(self):
        if self._h:
            _lib.cmd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop (cmd_server)."""

    def __init__(self, ctx: Context, max_batch: int = 4096, max_inflight: int = 2, queue_capacity: int = 1 << 20,
                 min_batch: int = 1, max_wait_us: float = 200.0, io_budget_bytes: int = 0):
        cfg = ServerConfig(max_batch, max_inflight, queue_capacity, min_batch, max_wait_us, io_budget_bytes)
        h = _P()
        _check(_lib.cmd_server_create(ctx.handle, _c.byref(cfg), _c.byref(h)), "cmd_server_create")
        self._h = h
        self.ctx = ctx

    def submit(self, key_id: int, ct: np.ndarray) -> int:
        """One decryption request; the ciphertext is the key's ct_bytes (shorter buffers are refused)."""
        ct = np.ascontiguousarray(ct, dtype=np.uint8)
        pid = self.ctx.key_param.get(int(key_id))
        if pid is not None and ct.nbytes < params(pid).ct_bytes:
            raise ValueError(f"ciphertext of {ct.nbytes} bytes, {params(pid).ct_bytes} needed")
        rid = _U64(0)
        _check(_lib.cmd_server_submit_batch(self._h, 1, _ptr(np.array([key_id], np.int32)), _ptr(ct), ct.nbytes,
                                            _c.byref(rid), None), "cmd_server_submit")
        return int(rid.value)

    def submit_batch(self, key_ids, ct: np.ndarray, ct_stride: Optional[int] = None,
                     mode: int = MODE_FUSED) -> np.ndarray:
        """Submits len(key_ids) requests (rows of ct) of `mode` (MODE_FUSED: error vectors,
        MODE_KEM: session keys); returns their ids."""
        keys = np.ascontiguousarray(key_ids, dtype=np.int32)
        ct = np.ascontiguousarray(ct, dtype=np.uint8)
        if ct_stride is None:
            ct_stride = ct.shape[1] if ct.ndim == 2 else ct.size // max(keys.size, 1)
        _check_rows(ct, keys.size, ct_stride, "ct")
        rids = np.zeros(keys.size, np.uint64)
        acc = _SZ(0)
        _check(_lib.cmd_server_submit_batch_mode(self._h, mode, keys.size, _ptr(keys), _ptr(ct), ct_stride,
                                                 _ptr(rids), _c.byref(acc)), "cmd_server_submit_batch_mode")
        return rids

    def wait_batch(self, rids: np.ndarray, e_stride: int = 1024, timeout_us: float = -1.0) -> np.ndarray:
        """Results of the requests rids (rows of e_stride bytes: error vectors or session keys)."""
        rids = np.ascontiguousarray(rids, dtype=np.uint64)
        e = np.zeros((rids.size, e_stride), np.uint8)
        _check(_lib.cmd_server_wait_batch(self._h, rids.size, _ptr(rids), _ptr(e), e_stride, timeout_us),
               "cmd_server_wait_batch")
        return e

    def try_submit(self, key_id: int, ct: np.ndarray):
        ct = np.ascontiguousarray(ct, dtype=np.uint8)
        rid = _U64(0)
        rc = _lib.cmd_server_submit_batch(self._h, 1, _ptr(np.array([key_id], np.int32)), _ptr(ct), ct.nbytes,
                                          _c.byref(rid), None)
        return rc, int(rid.value)

    def result_bytes(self, req: int) -> int:
        n = _SZ(0)
        _check(_lib.cmd_server_result_bytes(self._h, req, _c.byref(n)), "cmd_server_result_bytes")
        return int(n.value)

    def poll(self, req: int, e_bytes: int):
        e = np.zeros(e_bytes, np.uint8)
        rc = _lib.cmd_server_poll(self._h, req, _ptr(e), e.nbytes)
        if rc == CMD_ERR_PENDING:
            return None
        _check(rc, "cmd_server_poll")
        return e

    def wait(self, req: int, e_bytes: int, timeout_us: float = -1.0) -> np.ndarray:
        e = np.zeros(e_bytes, np.uint8)
        _check(_lib.cmd_server_wait(self._h, req, _ptr(e), e.nbytes, timeout_us), "cmd_server_wait")
        return e

    def latencies(self):
        """(total_us, queue_us) per retired request since the last reset_stats()."""
        n = _SZ(0)
        _check(_lib.cmd_server_latencies(self._h, None, None, 0, _c.byref(n)), "cmd_server_latencies")
        tot = np.zeros(n.value, np.float64)
        que = np.zeros(n.value, np.float64)
        _check(_lib.cmd_server_latencies(self._h, _ptr(tot), _ptr(que), n.value, _c.byref(n)),
               "cmd_server_latencies")
        return tot, que

    LAUNCH_DTYPE = np.dtype([("t_us", "<f8"), ("enqueue_us", "<f8"), ("gpu_us", "<f8"), ("seen_us", "<f8"),
                             ("retire_us", "<f8"), ("items", "<u4"), ("pad", "<u4")])

    def launch_log(self) -> np.ndarray:
        """Per-launch timing records since the last reset_stats() (cmd_server_launch)."""
        n = _SZ(0)
        _check(_lib.cmd_server_launch_log(self._h, None, 0, _c.byref(n)), "cmd_server_launch_log")
        out = np.zeros(n.value, self.LAUNCH_DTYPE)
        _check(_lib.cmd_server_launch_log(self._h, _ptr(out), n.value, _c.byref(n)), "cmd_server_launch_log")
        return out[:n.value]

    def stats(self) -> dict:
        st = ServerStats()
        _check(_lib.cmd_server_stats_get(self._h, _c.byref(st)), "cmd_server_stats_get")
        return st.as_dict()

    def reset_stats(self):
        _check(_lib.cmd_server_reset_stats(self._h), "cmd_server_reset_stats")

    def close(self):
        if self._h:
            _lib.cmd_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
