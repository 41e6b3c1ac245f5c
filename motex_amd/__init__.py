"""MI355X (gfx950) all-pairs motif extraction -- Python binding of the C ABI.

The product path is libmotex.so (hand-written HIP kernels + C++ runtime) reached through
the extern "C" interface declared in include/motex.h.  This module only marshals
arguments; it has no compute fallback and raises if the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MOTEX_LIB points at another build of the same library (e.g. an experimental variant).
LIB_PATH = os.environ.get("MOTEX_LIB", os.path.join(HERE, "lib", "libmotex.so"))
BIN_PATH = os.path.join(HERE, "bin", "motex")

ALPHA = {"DNA": 0, "PROT": 1, "USR": 2}

u32p = C.POINTER(C.c_uint32)


class MotexError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"motex error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("k", C.c_uint32), ("e", C.c_uint32), ("distance", C.c_uint32),
                ("alphabet", C.c_uint32), ("nb_gaps", C.c_uint32),
                ("gap_min", u32p), ("gap_max", u32p), ("box_len", u32p), ("box_err", u32p),
                ("long_mode", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("cells", C.c_double), ("kernel_ms", C.c_double), ("main_kernel_ms", C.c_double),
                ("work_items", C.c_uint64), ("pairs", C.c_uint64), ("skipped_pairs", C.c_uint64),
                ("tile_blocks", C.c_uint32), ("tile_rows", C.c_uint32),
                ("waves_per_cu", C.c_uint32), ("kernel_variant", C.c_uint32),
                ("graph_replay", C.c_uint32), ("symmetric", C.c_uint32),
                ("swept_cells", C.c_double), ("handed_pairs", C.c_uint64)]


class Tile(C.Structure):
    _fields_ = [("blocks", C.c_uint32), ("rows", C.c_uint32), ("waves_per_cu", C.c_uint32),
                ("candidates", C.c_uint32), ("kernel_ms", C.c_double)]


class Job(C.Structure):
    _fields_ = [("seqs", C.POINTER(C.c_char_p)), ("lens", u32p), ("num_seqs", C.c_uint32),
                ("params", Params), ("u", C.POINTER(u32p)), ("v", C.POINTER(u32p))]


# (name, restype, argtypes) of every symbol declared in include/motex.h
SIGNATURES = [
    ("motex_ctx_create", C.c_int, [C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]),
    ("motex_ctx_destroy", None, [C.c_void_p]),
    ("motex_last_error", C.c_char_p, []),
    ("motex_extract_all_pairs", C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), u32p, C.c_uint32,
                                          C.POINTER(Params), C.POINTER(u32p), C.POINTER(u32p)]),
    ("motex_extract_pair", C.c_int, [C.c_void_p, C.c_char_p, C.c_uint32, C.c_char_p, C.c_uint32,
                                     C.POINTER(Params), u32p, u32p]),
    ("motex_extract_cross", C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), u32p, C.c_uint32,
                                      C.POINTER(C.c_char_p), u32p, C.c_uint32, C.POINTER(Params),
                                      C.POINTER(u32p), C.POINTER(u32p)]),
    ("motex_extract_all_pairs_noalloc", C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), u32p, C.c_uint32,
                                                  C.POINTER(Params), C.POINTER(u32p), C.POINTER(u32p)]),
    ("motex_batch_create", C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), u32p, C.c_uint32,
                                     C.POINTER(Params), C.POINTER(C.c_void_p)]),
    ("motex_batch_run", C.c_int, [C.c_void_p]),
    ("motex_batch_run_plain", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    ("motex_batch_clear", C.c_int, [C.c_void_p]),
    ("motex_batch_download", C.c_int, [C.c_void_p, C.POINTER(u32p), C.POINTER(u32p)]),
    ("motex_batch_stats", C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    ("motex_batch_destroy", None, [C.c_void_p]),
    ("motex_autotune", C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), u32p, C.c_uint32,
                                 C.POINTER(Params), C.POINTER(Tile)]),
    ("motex_submit", C.c_int, [C.c_void_p, C.POINTER(Job), C.POINTER(C.c_uint64)]),
    ("motex_poll", C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int)]),
    ("motex_wait", C.c_int, [C.c_void_p, C.c_uint64]),
    ("motex_job_release", C.c_int, [C.c_void_p, C.c_uint64]),
    ("motex_job_times", C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]),
]

_lib = None


def lib():
    """Loads libmotex.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m motex_amd.build` first")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise MotexError(rc, lib().motex_last_error().decode(errors="replace"))


@dataclass
class MotexParams:
    k: int
    e: int
    distance: int = 0
    alphabet: str = "DNA"
    boxes: list = field(default_factory=list)   # [(gap_min, gap_max, box_len, box_err)]
    long_mode: bool = False

    def to_c(self):
        arrs = [np.ascontiguousarray(np.array([b[i] for b in self.boxes] or [0], np.uint32))
                for i in range(4)]
        p = Params(self.k, self.e, self.distance, ALPHA[self.alphabet], len(self.boxes),
                   *(a.ctypes.data_as(u32p) for a in arrs), int(self.long_mode))
        return p, arrs

    @property
    def nb_structs(self) -> int:
        n = 1
        for b in self.boxes:
            n *= b[1] - b[0] + 1
        return n


def _seq_args(seqs: list[bytes]):
    arr = (C.c_char_p * len(seqs))(*seqs)
    lens = np.ascontiguousarray(np.array([len(s) for s in seqs] or [0], np.uint32))
    return arr, lens


def uv_layout(seqs: list[bytes], nb_structs: int = 1):
    lens = np.array([len(s) for s in seqs], dtype=np.uint64) * np.uint64(nb_structs)
    off = np.zeros(len(seqs) + 1, dtype=np.uint64)
    off[1:] = np.cumsum(lens)
    return off


def _row_ptrs(flat: np.ndarray, off: np.ndarray):
    base = flat.ctypes.data
    return (u32p * (len(off) - 1))(*[C.cast(base + 4 * int(o), u32p) for o in off[:-1]])


class Context:
    """A device context (motex_ctx)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.h = C.c_void_p()
        _check(lib().motex_ctx_create(device, hbm_budget, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().motex_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def extract_all_pairs(self, seqs: list[bytes], params: MotexParams):
        """Batched persistent form.  Returns flat (u, v, offsets)."""
        off = uv_layout(seqs, params.nb_structs)
        u = np.zeros(max(int(off[-1]), 1), np.uint32)
        v = np.zeros(max(int(off[-1]), 1), np.uint32)
        sa, lens = _seq_args(seqs)
        p, keep = params.to_c()
        _check(lib().motex_extract_all_pairs(self.h, sa, lens.ctypes.data_as(u32p), len(seqs),
                                             C.byref(p), _row_ptrs(u, off), _row_ptrs(v, off)))
        return u[: int(off[-1])], v[: int(off[-1])], off

    def extract_cross(self, xs: list[bytes], ys: list[bytes], params: MotexParams):
        """Every x against every text y only (the -I re-search).  Returns flat (u, v, offsets)
        over the x sequences."""
        off = uv_layout(xs, params.nb_structs)
        u = np.zeros(max(int(off[-1]), 1), np.uint32)
        v = np.zeros(max(int(off[-1]), 1), np.uint32)
        xa, xl = _seq_args(xs)
        ya, yl = _seq_args(ys)
        p, keep = params.to_c()
        _check(lib().motex_extract_cross(self.h, xa, xl.ctypes.data_as(u32p), len(xs), ya,
                                         yl.ctypes.data_as(u32p), len(ys), C.byref(p),
                                         _row_ptrs(u, off), _row_ptrs(v, off)))
        return u[: int(off[-1])], v[: int(off[-1])], off

    def extract_all_pairs_noalloc(self, seqs: list[bytes], params: MotexParams):
        """Hamming all-pairs form that allocates no device memory (works in the library's
        static per-device scratch array).  Returns flat (u, v, offsets)."""
        off = uv_layout(seqs, params.nb_structs)
        u = np.zeros(max(int(off[-1]), 1), np.uint32)
        v = np.zeros(max(int(off[-1]), 1), np.uint32)
        sa, lens = _seq_args(seqs)
        p, keep = params.to_c()
        _check(lib().motex_extract_all_pairs_noalloc(self.h, sa, lens.ctypes.data_as(u32p), len(seqs),
                                                     C.byref(p), _row_ptrs(u, off), _row_ptrs(v, off)))
        return u[: int(off[-1])], v[: int(off[-1])], off

    def extract_pair(self, x: bytes, y: bytes, params: MotexParams, u=None, v=None):
        """Plain one-launch-per-pair form; accumulates into (and returns) u, v."""
        n = params.nb_structs * len(x)
        u = np.zeros(max(n, 1), np.uint32) if u is None else u
        v = np.zeros(max(n, 1), np.uint32) if v is None else v
        p, keep = params.to_c()
        _check(lib().motex_extract_pair(self.h, x, len(x), y, len(y), C.byref(p),
                                        u.ctypes.data_as(u32p), v.ctypes.data_as(u32p)))
        return u[:n], v[:n]

    def autotune(self, seqs: list[bytes], params: MotexParams) -> dict:
        """Times the tile shapes on this sample; later batches with the same key use the
        winner.  Returns the winning shape."""
        sa, lens = _seq_args(seqs)
        p, keep = params.to_c()
        t = Tile()
        _check(lib().motex_autotune(self.h, sa, lens.ctypes.data_as(u32p), len(seqs), C.byref(p),
                                    C.byref(t)))
        return {f: getattr(t, f) for f, _ in Tile._fields_}

    def batch(self, seqs: list[bytes], params: MotexParams) -> "Batch":
        return Batch(self, seqs, params)


class Batch:
    """Device-resident batch (inputs packed in HBM; results stay on the device)."""

    def __init__(self, ctx: Context, seqs: list[bytes], params: MotexParams):
        self.ctx, self.seqs, self.params = ctx, seqs, params
        self._sa, self._lens = _seq_args(seqs)
        p, self._keep = params.to_c()
        self.h = C.c_void_p()
        _check(lib().motex_batch_create(ctx.h, self._sa, self._lens.ctypes.data_as(u32p),
                                        len(seqs), C.byref(p), C.byref(self.h)))

    def run(self):
        _check(lib().motex_batch_run(self.h))
        return self.stats()

    def run_plain(self, first_x: int = 0, num_x: int | None = None):
        n = len(self.seqs) if num_x is None else num_x
        _check(lib().motex_batch_run_plain(self.h, first_x, n))
        return self.stats()

    def clear(self):
        _check(lib().motex_batch_clear(self.h))

    def download(self):
        off = uv_layout(self.seqs, self.params.nb_structs)
        u = np.zeros(max(int(off[-1]), 1), np.uint32)
        v = np.zeros(max(int(off[-1]), 1), np.uint32)
        _check(lib().motex_batch_download(self.h, _row_ptrs(u, off), _row_ptrs(v, off)))
        return u[: int(off[-1])], v[: int(off[-1])], off

    def stats(self) -> dict:
        s = Stats()
        _check(lib().motex_batch_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def close(self):
        if self.h:
            lib().motex_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
