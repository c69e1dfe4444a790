"""ctypes access to the CPU oracle (oracle/build/liborc.so) for the tests.

The oracle is the checker: tests call it to compute expected results and
compare them with the product's.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_LIB = os.path.join(ROOT, "oracle", "build", "liborc.so")
ORC_CLI = os.path.join(ROOT, "oracle", "build", "orc_cluster")

_orc = None


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(ORC_LIB):
            raise RuntimeError("oracle not built; run __graft_entry__.build()")
        L = C.CDLL(ORC_LIB)
        L.orc_kmap.restype = C.c_int
        L.orc_kmap.argtypes = [C.c_float]
        L.orc_cutoffs.restype = None
        L.orc_cutoffs.argtypes = [C.c_int, C.c_float, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_lcs.restype = C.c_int
        L.orc_lcs.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int]
        L.orc_decide_fast.restype = C.c_int
        L.orc_decide_fast.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_int64)]
        L.orc_summin.restype = C.c_longlong
        L.orc_summin.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.c_int]
        L.orc_shared_words_pass.restype = C.c_int
        L.orc_shared_words_pass.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_sort_order.restype = None
        L.orc_sort_order.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_cluster_file.restype = C.c_longlong
        L.orc_cluster_file.argtypes = [C.c_char_p, C.c_char_p, C.c_float, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                       C.c_void_p, C.c_longlong, C.c_longlong, C.c_void_p, C.c_int]
        L.orc_read_fasta.restype = C.c_longlong
        L.orc_read_fasta.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong,
                                     C.c_longlong, C.POINTER(C.c_longlong)]
        _orc = L
    return _orc


CODE = {ord(c): i for i, c in enumerate("ACGT")}


def codes(seq):
    """Raw bytes -> code array 0..4 (func.cpp:155-188)."""
    if isinstance(seq, str):
        seq = seq.encode()
    table = np.full(256, 4, dtype=np.uint8)
    for ch, v in zip(b"ACGTacgtUu", [0, 1, 2, 3, 0, 1, 2, 3, 3, 3]):
        table[ch] = v
    return table[np.frombuffer(seq, dtype=np.uint8)]


def lcs(a, b):
    ca, cb = codes(a), codes(b)
    return orc().orc_lcs(ca.tobytes(), len(ca), cb.tobytes(), len(cb))


def decide_fast(a, b, bc):
    ca, cb = codes(a), codes(b)
    cells = C.c_int64(0)
    r = orc().orc_decide_fast(ca.tobytes(), len(ca), cb.tobytes(), len(cb), bc, C.byref(cells))
    return bool(r), cells.value


def summin(a, b, k):
    ca, cb = codes(a), codes(b)
    return orc().orc_summin(ca.tobytes(), len(ca), cb.tobytes(), len(cb), k)


def shared_words_pass(a, b, k, wc, early):
    """The shared-word test of q = b against representative a at cutoff wc
    (early: with the exact deficit stop)."""
    ca, cb = codes(a), codes(b)
    return bool(orc().orc_shared_words_pass(ca.tobytes(), len(ca), cb.tobytes(), len(cb), k, wc, 1 if early else 0))


def cutoffs(L, t, k):
    bc, wc = C.c_int(), C.c_int()
    orc().orc_cutoffs(L, t, k, C.byref(bc), C.byref(wc))
    return bc.value, wc.value


def kmap(t):
    return orc().orc_kmap(t)


def sort_order(lens):
    lens = np.ascontiguousarray(lens, dtype=np.int64)
    perm = np.zeros(len(lens), dtype=np.int64)
    orc().orc_sort_order(lens.ctypes.data, len(lens), perm.ctypes.data)
    return perm


def cluster_file(path, t, k=0, threads=8, fast=True, out_path="", max_rounds=-1, drop=False):
    """Oracle run of the whole path on a FASTA file (drop: no composition or
    shared-word filter, the cuda-version's drop mode).

    Returns (cluster[] over sorted indices, perm, stats dict)."""
    n_cap = 1 << 22
    cl = np.zeros(n_cap, dtype=np.int32)
    perm = np.zeros(n_cap, dtype=np.int64)
    st = np.zeros(8, dtype=np.float64)
    n = orc().orc_cluster_file(os.fsencode(path), os.fsencode(out_path), t, k, threads, 1 if fast else 0,
                               cl.ctypes.data, perm.ctypes.data, n_cap, max_rounds, st.ctypes.data,
                               1 if drop else 0)
    if n < 0:
        raise RuntimeError(f"oracle failed on {path}: {n}")
    keys = ["reps", "pairs", "rej_comp", "rej_word", "aligned", "accepted", "seconds"]
    return cl[:n].copy(), perm[:n].copy(), dict(zip(keys, st[:7].tolist()))


def read_fasta(path):
    """Oracle record split: list of (name, data) byte strings."""
    tot = C.c_longlong()
    n = orc().orc_read_fasta(os.fsencode(path), None, None, None, None, 0, 0, C.byref(tot))
    if n < 0:
        raise RuntimeError("oracle could not read " + path)
    size = max(os.path.getsize(path), 1)
    names = C.create_string_buffer(size)
    data = C.create_string_buffer(size)
    noff = np.zeros(n + 1, dtype=np.int64)
    doff = np.zeros(n + 1, dtype=np.int64)
    orc().orc_read_fasta(os.fsencode(path), names, noff.ctypes.data, data, doff.ctypes.data, n, size, C.byref(tot))
    nb, db = names.raw, data.raw
    return [(nb[noff[i]:noff[i + 1]], db[doff[i]:doff[i + 1]]) for i in range(n)]
