"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: imported by tests/, the smoke entry and bench.py's cpu_baseline
leg as the checker.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_pair.argtypes = [C.c_int, C.c_char_p, C.c_uint32, C.c_char_p, C.c_uint32,
                               C.c_uint32, C.c_uint32, u32p, u32p]
        L.orc_hits.argtypes = [C.c_int, C.c_char_p, C.c_uint32, C.c_char_p, C.c_uint32,
                               C.c_uint32, C.c_uint32, u32p, C.c_uint64]
        L.orc_hits.restype = C.c_int64
        L.orc_struct_pair.argtypes = [C.c_int, C.c_char_p, C.c_uint32, C.c_char_p, C.c_uint32,
                                      C.c_uint32, C.c_uint32, C.c_uint32, u32p, u32p,
                                      C.c_uint32, u32p, u32p, u32p]
        L.orc_recurse.argtypes = [C.c_uint32, u32p, u32p, u32p]
        L.orc_recurse.restype = C.c_uint32
        L.orc_all_pairs.argtypes = [C.c_int, C.c_char_p, u64p, u32p, C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_uint32, u32p, u32p, u32p, u32p, u64p,
                                    u32p, u32p, C.c_int]
        L.orc_all_pairs.restype = C.c_int64
        L.orc_all_pairs_range.argtypes = [C.c_int, C.c_char_p, u64p, u32p, C.c_uint32, C.c_uint32,
                                          C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u32p,
                                          u32p, u32p, u32p, u64p, u32p, u32p, C.c_int]
        L.orc_all_pairs_range.restype = C.c_int64
        L.orc_bitminmax.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_bitminmax.restype = C.c_uint64
        L.orc_bitminmax_w.argtypes = [C.c_uint64] * 6 + [u64p]
        L.orc_write.argtypes = [C.c_void_p, C.c_char_p, u64p, u32p, C.c_uint32, u64p, u32p, u32p,
                                C.c_double, C.c_int, C.c_char_p, C.c_char_p]
        L.orc_write.restype = C.c_int64
        _lib = L
    return _lib


def _p32(a):
    return a.ctypes.data_as(u32p)


def _p64(a):
    return a.ctypes.data_as(u64p)


def _arr32(xs):
    return np.ascontiguousarray(np.asarray(xs if len(xs) else [0], dtype=np.uint32))


def pair(distance: int, x: bytes, y: bytes, k: int, e: int, u=None, v=None):
    """Per-pair operator; accumulates into (and returns) u, v of length len(x)."""
    m = len(x)
    u = np.zeros(m, np.uint32) if u is None else u
    v = np.zeros(m, np.uint32) if v is None else v
    rc = lib().orc_pair(distance, x, m, y, len(y), k, e, _p32(u), _p32(v))
    if rc < 0:
        raise ValueError(f"orc_pair failed ({rc})")
    return u, v


def hits(distance: int, x: bytes, y: bytes, k: int, e: int) -> np.ndarray:
    """FLASM tuples as an (n, 3) array of (pos_t, pos_x, err)."""
    cap = 1 << 16
    while True:
        out = np.zeros(3 * cap, np.uint32)
        cnt = lib().orc_hits(distance, x, len(x), y, len(y), k, e, _p32(out), cap)
        if cnt < 0:
            raise ValueError("orc_hits failed")
        if cnt <= cap:
            return out[: 3 * cnt].reshape(-1, 3)
        cap = int(cnt)


def recurse(gmin, gmax) -> np.ndarray:
    g0, g1 = _arr32(gmin), _arr32(gmax)
    ng = len(gmin)
    ns = lib().orc_recurse(ng, _p32(g0), _p32(g1), None)
    S = np.zeros(max(1, ns * ng), np.uint32)
    lib().orc_recurse(ng, _p32(g0), _p32(g1), _p32(S))
    return S[: ns * ng].reshape(ns, ng)


def struct_pair(distance, x, y, k, e, boxes, u=None, v=None):
    """boxes: list of (gmin, gmax, len, err).  u, v have nb_structs*len(x) entries."""
    gmin = [b[0] for b in boxes]; gmax = [b[1] for b in boxes]
    S = recurse(gmin, gmax)
    ns = S.shape[0]
    m = len(x)
    u = np.zeros(ns * m, np.uint32) if u is None else u
    v = np.zeros(ns * m, np.uint32) if v is None else v
    bl = _arr32([b[2] for b in boxes]); be = _arr32([b[3] for b in boxes])
    S = np.ascontiguousarray(S.astype(np.uint32).ravel() if S.size else np.zeros(1, np.uint32))
    rc = lib().orc_struct_pair(distance, x, m, y, len(y), k, e, len(boxes), _p32(bl), _p32(be),
                               ns, _p32(S), _p32(u), _p32(v))
    if rc < 0:
        raise ValueError("orc_struct_pair failed")
    return u, v


def pack(seqs: list[bytes]):
    lens = np.array([len(s) for s in seqs], dtype=np.uint32)
    off = np.zeros(len(seqs), dtype=np.uint64)
    if len(seqs) > 1:
        off[1:] = np.cumsum(lens[:-1].astype(np.uint64))
    return b"".join(seqs), off, lens


def all_pairs(seqs: list[bytes], distance: int, k: int, e: int, boxes=(), threads: int = 0,
              x_range=None):
    """Returns (u, v, uv_off, skipped) with u, v flat (nb_structs*len per sequence).
    x_range=(x0, x1) restricts the x sequences (every y is still searched)."""
    blob, off, lens = pack(seqs)
    gmin = _arr32([b[0] for b in boxes]); gmax = _arr32([b[1] for b in boxes])
    bl = _arr32([b[2] for b in boxes]); be = _arr32([b[3] for b in boxes])
    ns = 1
    for b in boxes:
        ns *= b[1] - b[0] + 1
    sizes = lens.astype(np.uint64) * np.uint64(ns)
    uv_off = np.zeros(len(seqs), np.uint64)
    if len(seqs) > 1:
        uv_off[1:] = np.cumsum(sizes[:-1])
    tot = int(sizes.sum())
    u = np.zeros(max(tot, 1), np.uint32)
    v = np.zeros(max(tot, 1), np.uint32)
    x0, x1 = x_range if x_range else (0, len(seqs))
    sk = lib().orc_all_pairs_range(distance, blob, _p64(off), _p32(lens), len(seqs), x0, x1, k, e,
                                   len(boxes), _p32(gmin), _p32(gmax), _p32(bl), _p32(be),
                                   _p64(uv_off), _p32(u), _p32(v), threads)
    if sk < 0:
        raise ValueError("orc_all_pairs failed")
    return u[:tot], v[:tot], uv_off, int(sk)


class _SW(C.Structure):
    _fields_ = [("alphabet", C.c_char_p), ("input_filename", C.c_char_p),
                ("output_filename", C.c_char_p), ("boxes_filename", C.c_char_p),
                ("q", C.c_uint32), ("Q", C.c_uint32), ("k", C.c_uint32), ("d", C.c_uint32),
                ("e", C.c_uint32), ("n", C.c_uint32), ("N", C.c_uint32), ("nb_gaps", C.c_uint32),
                ("gmin", u32p), ("gmax", u32p), ("blens", u32p), ("berrs", u32p),
                ("total_length", C.c_uint32)]


def write(path, seqs, u, v, uv_off, *, alphabet, d, k, e, q, Q=100, n=1, N=10000, boxes=(),
          input_filename="in.fa", boxes_filename="boxes.txt", smile_path=None, secs=0.0, P=1):
    """Write .motex (and SMILE) output with the oracle's writers; returns #motifs."""
    blob, off, lens = pack(seqs)
    g = [_arr32([b[i] for b in boxes]) for i in range(4)]
    sw = _SW(alphabet.encode(), input_filename.encode(), path.encode(), boxes_filename.encode(),
             q, Q, k, d, e, n, N, len(boxes), _p32(g[0]), _p32(g[1]), _p32(g[2]), _p32(g[3]),
             int(lens.sum()))
    u = np.ascontiguousarray(u, np.uint32); v = np.ascontiguousarray(v, np.uint32)
    return lib().orc_write(C.byref(sw), blob, _p64(off), _p32(lens), len(seqs),
                           _p64(np.ascontiguousarray(uv_off, np.uint64)), _p32(u), _p32(v),
                           secs, P, path.encode(), smile_path.encode() if smile_path else None)
