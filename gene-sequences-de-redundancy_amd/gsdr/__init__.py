"""ctypes binding of libgsdr.so, the MI355X de-redundancy engine (include/gsdr.h).

The library is built in-tree (``gene-sequences-de-redundancy_amd/build``) by
``__graft_entry__.build()`` or ``make``. Loading fails loudly when it is
missing: there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD_DIR = os.path.join(PKG_DIR, "build")
LIB_PATH = os.path.join(BUILD_DIR, "libgsdr.so")
CLI_PATH = os.path.join(BUILD_DIR, "cluster")

MODE_PLAIN = 0
MODE_BATCHED = 1
MAX_GAPFREE = 32767

OK, E_INVALID, E_IO, E_INPUT, E_DEVICE, E_NOMEM, E_LIMIT, E_BUSY, E_NOTFOUND = range(9)


class Params(C.Structure):
    _fields_ = [
        ("threshold", C.c_float),
        ("word_length", C.c_int32),
        ("mode", C.c_int32),
        ("spec_reps", C.c_int32),
        ("timing", C.c_int32),
        ("tune", C.c_int32),
        ("drop_filters", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("n_seqs", C.c_int64),
        ("n_bases", C.c_int64),
        ("reps", C.c_int64),
        ("rounds", C.c_int64),
        ("launches", C.c_int64),
        ("pairs", C.c_int64),
        ("rej_length", C.c_int64),
        ("rej_comp", C.c_int64),
        ("rej_word", C.c_int64),
        ("aligned", C.c_int64),
        ("rej_align", C.c_int64),
        ("accepted", C.c_int64),
        ("cells_band", C.c_double),
        ("cells_exec", C.c_double),
        ("cells_full", C.c_double),
        ("ms_total", C.c_double),
        ("ms_filter", C.c_double),
        ("ms_align", C.c_double),
        ("filter_launches", C.c_int64),
        ("align_launches", C.c_int64),
        ("filter_bytes", C.c_double),
        ("spec_wasted", C.c_double),
        ("cells_decided", C.c_double),
        ("rej_pigeon", C.c_double),
        ("pigeon_checked", C.c_double),
        ("screened", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [
        ("swap_ratio", C.c_int32),
        ("spec_max", C.c_int32),
        ("filter_blocks", C.c_int32),
        ("align_blocks", C.c_int32),
        ("trials", C.c_int32),
        ("fork_shapes", C.c_int32),
        ("list_factor", C.c_int32),
        ("pigeon", C.c_int32),
        ("screen", C.c_int32),
        ("ms_default", C.c_double),
        ("ms_tuned", C.c_double),
        ("shape_mask", C.c_int32 * 4),
        ("shape_bias", C.c_int32 * 4),
        ("win_period", C.c_int32),
        ("split_rows", C.c_int32),
    ]

    def __init__(self, **kw):
        # the signature screen is on unless a caller turns it off
        kw.setdefault("screen", 1)
        # grids and shape concurrency chosen by the call (the library default)
        kw.setdefault("align_blocks", -1)
        kw.setdefault("filter_blocks", -1)
        kw.setdefault("fork_shapes", -1)
        kw.setdefault("split_rows", -1)
        super().__init__(**kw)

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if name in ("shape_mask", "shape_bias") else v
        return out


class Options(C.Structure):
    """gsdr_options: the reference's Option (func.h:11-16) plus long options."""
    _fields_ = [("input", C.c_char_p), ("output", C.c_char_p), ("threshold", C.c_float), ("word_length", C.c_int32),
                ("mode", C.c_int32), ("spec_reps", C.c_int32), ("device", C.c_int32), ("drop_filters", C.c_int32),
                ("membership", C.c_char_p), ("report", C.c_char_p)]


# every entry point of include/gsdr.h: (name, restype, argtypes)
_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
SIGNATURES = [
    ("gsdr_parse_options", C.c_int, [C.c_int, C.POINTER(C.c_char_p), C.POINTER(Options)]),
    ("gsdr_cutoffs", C.c_int, [C.c_float, C.c_int32, C.POINTER(C.c_int32), C.c_int64, C.POINTER(C.c_int32),
                               C.POINTER(C.c_int32)]),
    ("gsdr_records_read_fasta", C.c_int, [C.c_char_p, _PP]),
    ("gsdr_records_read_fasta_sorted", C.c_int, [C.c_char_p, _PP]),
    ("gsdr_records_from_buffers", C.c_int, [C.c_char_p, C.POINTER(C.c_int64), C.c_char_p,
                                            C.POINTER(C.c_int64), C.c_int64, _PP]),
    ("gsdr_records_generate", C.c_int, [C.c_int32, C.c_uint64, C.c_int64, C.c_double, _PP]),
    ("gsdr_records_write_fasta", C.c_int, [_P, C.c_char_p, C.c_int32]),
    ("gsdr_records_sort", C.c_int, [_P, C.POINTER(C.c_int64)]),
    ("gsdr_records_count", C.c_int64, [_P]),
    ("gsdr_records_bases", C.c_int64, [_P]),
    ("gsdr_records_get", C.c_int, [_P, C.c_int64, C.POINTER(C.c_char_p), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]),
    ("gsdr_records_write_result", C.c_int, [_P, C.POINTER(C.c_int32), C.c_char_p, C.POINTER(C.c_int64)]),
    ("gsdr_records_write_membership", C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.c_char_p]),
    ("gsdr_records_free", None, [_P]),
    ("gsdr_open", C.c_int, [C.c_int32, _PP]),
    ("gsdr_close", None, [_P]),
    ("gsdr_device_info", C.c_int, [_P, C.c_char_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    ("gsdr_load", C.c_int, [_P, _P, C.c_int32, _PP]),
    ("gsdr_store_free", None, [_P]),
    ("gsdr_store_footprint", C.c_int64, [_P]),
    ("gsdr_cluster", C.c_int, [_P, _P, C.POINTER(Params), C.POINTER(C.c_int32), C.POINTER(Stats)]),
    ("gsdr_cluster_noalloc", C.c_int, [_P, _P, C.POINTER(Params), C.POINTER(C.c_int32), C.POINTER(Stats)]),
    ("gsdr_cluster_jobs", C.c_int, [_P, C.c_int32, C.POINTER(C.c_void_p), C.POINTER(Params),
                                    C.POINTER(C.POINTER(C.c_int32)), C.POINTER(Stats)]),
    ("gsdr_autotune", C.c_int, [_P, _P, C.POINTER(Params), C.c_int32, C.POINTER(Tuning)]),
    ("gsdr_set_tuning", C.c_int, [_P, C.POINTER(Tuning)]),
    ("gsdr_get_tuning", C.c_int, [_P, C.POINTER(Tuning)]),
    ("gsdr_align_pairs", C.c_int, [_P, C.c_char_p, C.POINTER(C.c_int64), C.c_int32, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_double)]),
    ("gsdr_serve_open", C.c_int, [_P, C.c_int64, C.c_int32, _PP]),
    ("gsdr_serve_submit", C.c_int, [_P, _P, C.POINTER(Params), C.POINTER(C.c_int64)]),
    ("gsdr_serve_step", C.c_int, [_P, C.POINTER(C.c_int32)]),
    ("gsdr_serve_poll", C.c_int, [_P, C.c_int64, C.POINTER(C.c_int32), C.POINTER(Stats), C.POINTER(C.c_int32)]),
    ("gsdr_serve_close", None, [_P]),
    ("gsdr_strerror", C.c_char_p, [C.c_int32]),
    ("gsdr_last_error", C.c_char_p, []),
]

_lib = None


def lib():
    """Loads libgsdr.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libgsdr.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class GsdrError(RuntimeError):
    def __init__(self, status, what):
        L = lib()
        msg = L.gsdr_last_error().decode(errors="replace")
        super().__init__(f"{what}: {L.gsdr_strerror(status).decode()} ({status}): {msg}")
        self.status = status


def _check(st, what):
    if st != OK:
        raise GsdrError(st, what)


def _i64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _i32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Records:
    """FASTA records (names + raw data) owned by the library."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    @staticmethod
    def read_fasta(path, length_order=False):
        """File order, or with length_order the reference's length order
        (what sort() would give); sort() then only reports the permutation."""
        h = C.c_void_p()
        fn = lib().gsdr_records_read_fasta_sorted if length_order else lib().gsdr_records_read_fasta
        _check(fn(os.fsencode(path), C.byref(h)), "read_fasta")
        return Records(h)

    @staticmethod
    def generate(law, seed, n=0, median=0.0):
        h = C.c_void_p()
        _check(lib().gsdr_records_generate(law, seed, n, median, C.byref(h)), "generate")
        return Records(h)

    @staticmethod
    def from_lists(datas, names=None):
        datas = [d if isinstance(d, bytes) else d.encode() for d in datas]
        if names is None:
            names = [b">s%d" % i for i in range(len(datas))]
        names = [n if isinstance(n, bytes) else n.encode() for n in names]
        doff = np.zeros(len(datas) + 1, dtype=np.int64)
        noff = np.zeros(len(names) + 1, dtype=np.int64)
        doff[1:] = np.cumsum([len(d) for d in datas])
        noff[1:] = np.cumsum([len(n) for n in names])
        h = C.c_void_p()
        _check(lib().gsdr_records_from_buffers(b"".join(names), _i64p(noff), b"".join(datas), _i64p(doff),
                                               len(datas), C.byref(h)), "from_buffers")
        return Records(h)

    def __len__(self):
        return lib().gsdr_records_count(self.h)

    @property
    def bases(self):
        return lib().gsdr_records_bases(self.h)

    def get(self, i):
        nm = C.c_char_p()
        nl = C.c_int64()
        dp = C.c_void_p()
        dl = C.c_int64()
        _check(lib().gsdr_records_get(self.h, i, C.byref(nm), C.byref(nl), C.byref(dp), C.byref(dl)), "get")
        name = C.string_at(nm, nl.value) if nl.value else b""
        data = C.string_at(dp, dl.value) if dl.value else b""
        return name, data

    def sort(self):
        perm = np.zeros(len(self), dtype=np.int64)
        _check(lib().gsdr_records_sort(self.h, _i64p(perm)), "sort")
        return perm

    def write_fasta(self, path, width=80):
        _check(lib().gsdr_records_write_fasta(self.h, os.fsencode(path), width), "write_fasta")

    def write_result(self, cluster, path):
        reps = C.c_int64()
        cl = np.ascontiguousarray(cluster, dtype=np.int32)
        _check(lib().gsdr_records_write_result(self.h, _i32p(cl), os.fsencode(path), C.byref(reps)), "write_result")
        return reps.value

    def write_membership(self, perm, cluster, path):
        cl = np.ascontiguousarray(cluster, dtype=np.int32)
        pp = np.ascontiguousarray(perm, dtype=np.int64)
        _check(lib().gsdr_records_write_membership(self.h, _i64p(pp), _i32p(cl), os.fsencode(path)), "membership")

    def free(self):
        if self.h:
            lib().gsdr_records_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def parse_options(argv):
    """checkOption (func.cpp:6-57) through the C ABI: argv without the
    program name. Returns an Options copy, or raises GsdrError with the
    reference's usage text."""
    args = [b"cluster"] + [a if isinstance(a, bytes) else a.encode() for a in argv]
    arr = (C.c_char_p * len(args))(*args)
    o = Options()
    _check(lib().gsdr_parse_options(len(args), arr, C.byref(o)), "parse_options")
    return {"input": o.input, "output": o.output, "threshold": o.threshold, "word_length": o.word_length,
            "mode": o.mode, "spec_reps": o.spec_reps, "device": o.device, "drop_filters": o.drop_filters,
            "membership": o.membership, "report": o.report}


def cutoffs(raw_lengths, threshold, word_length=0):
    """createCutoff (func.cpp:339-374): (bc, wc) arrays for raw lengths."""
    L = np.ascontiguousarray(raw_lengths, dtype=np.int32)
    bc = np.zeros(len(L), dtype=np.int32)
    wc = np.zeros(len(L), dtype=np.int32)
    _check(lib().gsdr_cutoffs(threshold, word_length, _i32p(L), len(L), _i32p(bc), _i32p(wc)), "cutoffs")
    return bc, wc


def auto_word_length(t):
    """func.cpp:36-45 (float threshold against double literals)."""
    t = float(np.float32(t))
    return 4 if t < 0.88 else 5 if t < 0.94 else 6 if t < 0.97 else 7


class Engine:
    """A device context on one GPU."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        _check(lib().gsdr_open(device, C.byref(self.h)), "open")

    def info(self):
        name = C.create_string_buffer(256)
        cu = C.c_int32()
        hbm = C.c_int64()
        _check(lib().gsdr_device_info(self.h, name, 256, C.byref(cu), C.byref(hbm)), "device_info")
        return {"name": name.value.decode(), "cu_count": cu.value, "hbm_bytes": hbm.value}

    def load(self, sorted_records, word_length):
        h = C.c_void_p()
        _check(lib().gsdr_load(self.h, sorted_records.h, word_length, C.byref(h)), "load")
        st = Store(self, h, len(sorted_records))
        st.k = word_length
        return st

    def align_pairs(self, datas, pairs, bc, shape=-1):
        """LCS(text, query) > bc per (text, query) pair, decided on the device."""
        datas = [d if isinstance(d, bytes) else d.encode() for d in datas]
        doff = np.zeros(len(datas) + 1, dtype=np.int64)
        doff[1:] = np.cumsum([len(d) for d in datas])
        pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        bcs = np.ascontiguousarray(np.asarray(bc, dtype=np.int32))
        out = np.zeros(len(pr), dtype=np.int32)
        cells = C.c_double()
        _check(lib().gsdr_align_pairs(self.h, b"".join(datas), _i64p(doff), len(datas), _i32p(pr), _i32p(bcs),
                                      len(pr), shape, _i32p(out), C.byref(cells)), "align_pairs")
        return out.astype(bool), cells.value

    def cluster_jobs(self, stores, thresholds, mode=MODE_BATCHED, spec_reps=0, drop_filters=None, tune=False):
        """Clusters several stores in shared launches; drop_filters: one flag
        per job (the drop mode), or None; tune: the context's tuned knobs."""
        n = len(stores)
        params = (Params * n)()
        for i, t in enumerate(thresholds):
            params[i].threshold = t
            params[i].word_length = stores[i].k if hasattr(stores[i], "k") else 0
            params[i].mode = mode
            params[i].spec_reps = spec_reps
            params[i].drop_filters = 1 if drop_filters is not None and drop_filters[i] else 0
            params[i].tune = 1 if tune else 0
        outs = [np.zeros(s.n, dtype=np.int32) for s in stores]
        handles = (C.c_void_p * n)(*[s.h.value for s in stores])
        ptrs = (C.POINTER(C.c_int32) * n)(*[_i32p(o) for o in outs])
        st = Stats()
        _check(lib().gsdr_cluster_jobs(self.h, n, handles, params, ptrs, C.byref(st)), "cluster_jobs")
        return outs, st

    def set_tuning(self, tuning=None):
        """Sets the knobs of calls with tune=True (None: back to the defaults)."""
        _check(lib().gsdr_set_tuning(self.h, C.byref(tuning) if tuning is not None else None), "set_tuning")

    def tuning(self):
        t = Tuning()
        _check(lib().gsdr_get_tuning(self.h, C.byref(t)), "get_tuning")
        return t

    def close(self):
        if self.h:
            lib().gsdr_close(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Store:
    """Sequences resident in HBM for one word length."""

    def __init__(self, engine, handle, n):
        self.engine = engine
        self.h = handle
        self.n = n
        self.k = None

    @property
    def footprint(self):
        return lib().gsdr_store_footprint(self.h)

    def cluster(self, threshold, word_length=0, mode=MODE_PLAIN, spec_reps=0, timing=False, tune=False,
                drop_filters=False, noalloc=False):
        """gsdr_cluster, or with noalloc gsdr_cluster_noalloc (the launch
        workspace in the library's fixed device array, no allocation)."""
        p = Params()
        p.drop_filters = 1 if drop_filters else 0
        p.threshold = threshold
        p.word_length = word_length
        p.mode = mode
        p.spec_reps = spec_reps
        p.timing = 1 if timing else 0
        p.tune = 1 if tune else 0
        out = np.zeros(self.n, dtype=np.int32)
        st = Stats()
        fn = lib().gsdr_cluster_noalloc if noalloc else lib().gsdr_cluster
        _check(fn(self.engine.h, self.h, C.byref(p), _i32p(out), C.byref(st)), "cluster")
        return out, st

    def autotune(self, threshold, word_length=0, spec_reps=0, max_trials=0):
        """Times batched clusterings of this store under different launch
        knobs and keeps the fastest in the engine (used with tune=True)."""
        p = Params()
        p.threshold = threshold
        p.word_length = word_length
        p.mode = MODE_BATCHED
        p.spec_reps = spec_reps
        t = Tuning()
        _check(lib().gsdr_autotune(self.engine.h, self.h, C.byref(p), max_trials, C.byref(t)), "autotune")
        return t

    def free(self):
        if self.h:
            lib().gsdr_store_free(self.h)
            self.h = None


def cluster_records(engine, sorted_records, threshold, word_length=0, mode=MODE_PLAIN, **kw):
    """Loads sorted records, clusters them and frees the store."""
    k = word_length or auto_word_length(threshold)
    store = engine.load(sorted_records, k)
    store.k = k
    try:
        return store.cluster(threshold, k, mode=mode, **kw)
    finally:
        store.free()
