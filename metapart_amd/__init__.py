"""metapart_amd: MI355X implementation of MetaPrep's read-partitioning path.

Python view of the C ABI in include/metapart_amd.h (ctypes; no torch types
cross the boundary).  The drop-in tools are the C programs in
metapart_amd/bin; this module is what tests, the benchmark and the batched /
served drivers use.  The library must be built (python -m metapart_amd.build);
nothing here falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as ct
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmetapart_amd.so")
SYNTH_PATH = os.path.join(PKG, "libmpsynth.so")
BIN = os.path.join(PKG, "bin")

NBUCKETS = 1 << 20
NPART = 16
FILEINDEX_BYTES = 4194344

ERRORS = {0: "ok", -1: "EINVAL", -2: "ENOMEM", -3: "EPLAN", -4: "ECOUNT", -5: "EDEVICE", -6: "EIO", -7: "ESTATE"}


class MPError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class Cfg(ct.Structure):
    _fields_ = [("device", ct.c_int), ("k", ct.c_int), ("min3", ct.c_int), ("hbm_budget", ct.c_uint64),
                ("sort_digit_bits", ct.c_int), ("sort_tile", ct.c_int), ("sort_graph", ct.c_int)]


class ScanStats(ct.Structure):
    _fields_ = [("tuples", ct.c_uint64), ("bad_labels", ct.c_uint64), ("zeroed_runs", ct.c_uint64),
                ("evaluated_runs", ct.c_uint64), ("quirk_segments", ct.c_int), ("sort_passes", ct.c_int),
                ("ms_enumerate", ct.c_float), ("ms_sort", ct.c_float), ("ms_filter", ct.c_float),
                ("ms_union", ct.c_float), ("ms_sort_pass_avg", ct.c_float), ("ms_sort_hist", ct.c_float),
                ("ms_sort_local", ct.c_float), ("sort_long_groups", ct.c_uint64), ("ms_sort_gen", ct.c_float),
                ("parts", ct.c_int), ("sort_local_moved", ct.c_uint64), ("ms_sort_count", ct.c_float),
                ("sort_graph", ct.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class FinishStats(ct.Structure):
    _fields_ = [("pair_unions", ct.c_uint64), ("pair_candidates", ct.c_uint64), ("components", ct.c_uint64),
                ("largest", ct.c_uint64), ("ms_device", ct.c_float), ("ms_host", ct.c_float)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SampleResult(ct.Structure):
    _fields_ = [("reads", ct.c_uint64), ("tuples", ct.c_uint64), ("components", ct.c_uint64),
                ("largest", ct.c_uint64), ("W", ct.c_uint64 * (NPART + 1))]

    def as_dict(self):
        return {"reads": self.reads, "tuples": self.tuples, "components": self.components, "largest": self.largest,
                "W": np.array(self.W[:], dtype=np.uint64)}


class ServerCfg(ct.Structure):
    _fields_ = [("device", ct.c_int), ("k", ct.c_int), ("engines", ct.c_int), ("max_jobs_per_batch", ct.c_int),
                ("hbm_budget", ct.c_uint64), ("min3", ct.c_int)]


class JobResult(ct.Structure):
    _fields_ = [("job", ct.c_uint64), ("status", ct.c_int), ("batch_jobs", ct.c_int), ("reads", ct.c_uint64),
                ("tuples", ct.c_uint64), ("components", ct.c_uint64), ("largest", ct.c_uint64),
                ("W", ct.c_uint64 * (NPART + 1)), ("t_submit", ct.c_double), ("t_start", ct.c_double),
                ("t_done", ct.c_double)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "W"}
        d["W"] = np.array(self.W[:], dtype=np.uint64)
        return d


_lib = None
_synth = None
P64 = ct.POINTER(ct.c_uint64)
MAX_SAMPLES = 256


def lib():
    """The loaded C library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmetapart_amd.so is not built: run python -m metapart_amd.build")
        L = ct.CDLL(LIB_PATH)
        vp, i, u64, cp = ct.c_void_p, ct.c_int, ct.c_uint64, ct.c_char_p
        sig = {
            "mp_index_fastq": (i, [i, i, i, i, ct.POINTER(cp), i]),
            "mp_index_buffer": (i, [vp, u64, i, i, P64, ct.POINTER(i)]),
            "mp_plan_ranges": (i, [P64, i, i, P64, P64]),
            "mp_ctx_create": (i, [ct.POINTER(Cfg), ct.POINTER(vp)]),
            "mp_ctx_destroy": (i, [vp]),
            "mp_last_error": (cp, [vp]),
            "mp_ctx_device_bytes": (u64, [vp]),
            "mp_count_chunks": (i, [vp, vp, u64, P64, i, vp, P64, P64]),
            "mp_count_loaded": (i, [vp, vp, P64, P64, P64]),
            "mp_chunk_first_ids": (i, [vp, P64]),
            "mp_load_reads": (i, [vp, vp, u64, P64, P64, i]),
            "mp_load_reads_device": (i, [vp, vp, u64, P64, P64, i]),
            "mp_scan": (i, [vp, i, u64, u64, P64, i, u64, ct.POINTER(ScanStats)]),
            "mp_finish": (i, [vp, P64, ct.POINTER(FinishStats)]),
            "mp_get_parent": (i, [vp, P64]),
            "mp_get_bad": (i, [vp, vp]),
            "mp_dump_tuples": (i, [vp, vp, u64, P64]),
            "mp_write_partitions": (i, [vp, cp, i, P64, ct.POINTER(ct.c_int64), P64]),
            "mp_batch_load": (i, [vp, i, ct.POINTER(vp), P64]),
            "mp_batch_load_device": (i, [vp, vp, u64, i, P64]),
            "mp_batch_partition": (i, [vp, ct.POINTER(SampleResult), ct.POINTER(ScanStats), ct.POINTER(FinishStats)]),
            "mp_batch_get_parent": (i, [vp, i, vp]),
            "mp_batch_write": (i, [vp, cp, P64]),
            "mp_server_create": (i, [ct.POINTER(ServerCfg), ct.POINTER(vp)]),
            "mp_server_submit": (i, [vp, vp, u64, i, P64]),
            "mp_server_wait": (i, [vp, u64, ct.POINTER(JobResult), vp, u64]),
            "mp_server_clock": (ct.c_double, [vp]),
            "mp_server_destroy": (i, [vp]),
            "mp_count_chunks_noalloc": (i, [i, i, i, vp, u64, P64, i, vp, P64, P64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


# exported symbol list (what include/metapart_amd.h declares)
SYMBOLS = ["mp_index_fastq", "mp_index_buffer", "mp_plan_ranges", "mp_ctx_create", "mp_ctx_destroy",
           "mp_last_error", "mp_ctx_device_bytes", "mp_count_chunks", "mp_count_loaded", "mp_chunk_first_ids",
           "mp_load_reads", "mp_load_reads_device", "mp_scan", "mp_finish", "mp_get_parent", "mp_get_bad", "mp_dump_tuples",
           "mp_write_partitions", "mp_batch_load", "mp_batch_load_device", "mp_batch_partition", "mp_batch_get_parent",
           "mp_batch_write", "mp_server_create", "mp_server_submit", "mp_server_wait", "mp_server_clock",
           "mp_server_destroy", "mp_count_chunks_noalloc"]


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def _p(a: np.ndarray):
    return a.ctypes.data_as(P64)


# ---------------------------------------------------------------- host-side functions
def index_buffer(buf, num_chunks: int, interleaved: bool = False) -> np.ndarray:
    """Chunk byte offsets of one in-memory FASTQ (genFastqIdx rules)."""
    ptr, n = _buf_ptr(buf)
    out = np.zeros(num_chunks + 1, dtype=np.uint64)
    nc = ct.c_int(0)
    rc = lib().mp_index_buffer(ptr, n, num_chunks, 1 if interleaved else 0, _p(out), ct.byref(nc))
    if rc:
        raise MPError(rc, "mp_index_buffer")
    return out[: nc.value + 1].copy()


def plan_ranges(hist: np.ndarray, S: int, T: int):
    hist = _u64(hist)
    scan = np.zeros(S + 1, dtype=np.uint64)
    thread = np.zeros(S * T + 1, dtype=np.uint64)
    rc = lib().mp_plan_ranges(_p(hist), S, T, _p(scan), _p(thread))
    if rc:
        raise MPError(rc, "mp_plan_ranges(S=%d, T=%d)" % (S, T))
    return scan, thread


def _buf_ptr(buf):
    if isinstance(buf, SynthBuffer):
        return buf.ptr, buf.nbytes
    if isinstance(buf, (bytes, bytearray)):
        arr = np.frombuffer(buf, dtype=np.uint8)
        return arr.ctypes.data, len(buf)
    arr = np.ascontiguousarray(buf, dtype=np.uint8)
    return arr.ctypes.data, arr.nbytes


# ---------------------------------------------------------------- device context
class Context:
    """One device context (mp_ctx)."""

    def __init__(self, k: int, min3: bool = False, device: int = 0, hbm_budget: int = 0, sort_tile: int = 0,
                 sort_graph: bool = False, sort_digit_bits: int = 0):
        self.k = k
        self.min3 = bool(min3)
        cfg = Cfg(device=device, k=k, min3=1 if min3 else 0, hbm_budget=hbm_budget, sort_digit_bits=sort_digit_bits,
                  sort_tile=sort_tile, sort_graph=1 if sort_graph else 0)
        h = ct.c_void_p()
        rc = lib().mp_ctx_create(ct.byref(cfg), ct.byref(h))
        if rc:
            raise MPError(rc, "mp_ctx_create (no usable GPU?)")
        self._h = h
        self.N = 0
        self.C = 0

    def close(self):
        if self._h:
            lib().mp_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc:
            raise MPError(rc, "%s: %s" % (what, lib().mp_last_error(self._h).decode(errors="replace")))

    @property
    def device_bytes(self) -> int:
        return int(lib().mp_ctx_device_bytes(self._h))

    def load(self, buf, chunk_ofs, first_ids=None):
        """Upload host FASTQ bytes; first_ids None lets the device count them."""
        cofs = _u64(chunk_ofs)
        ptr, n = _buf_ptr(buf)
        fid = _u64(first_ids) if first_ids is not None else None
        self._check(lib().mp_load_reads(self._h, ptr, n, _p(cofs), _p(fid) if fid is not None else None,
                                        len(cofs) - 1), "mp_load_reads")
        self.C = len(cofs) - 1
        self.N = int(self.first_ids()[-1])

    def load_device(self, dev_ptr: int, nbytes: int, chunk_ofs, first_ids=None):
        """FASTQ bytes already in device memory (e.g. a torch uint8 tensor's data_ptr())."""
        cofs = _u64(chunk_ofs)
        fid = _u64(first_ids) if first_ids is not None else None
        self._check(lib().mp_load_reads_device(self._h, ct.c_void_p(dev_ptr), nbytes, _p(cofs),
                                               _p(fid) if fid is not None else None, len(cofs) - 1),
                    "mp_load_reads_device")
        self.C = len(cofs) - 1
        self.N = int(self.first_ids()[-1])

    def first_ids(self) -> np.ndarray:
        out = np.zeros(self.C + 1, dtype=np.uint64)
        self._check(lib().mp_chunk_first_ids(self._h, _p(out)), "mp_chunk_first_ids")
        return out

    def count(self, want_counts: bool = False):
        """(hist prefix [2^20+1], reads per chunk, total k-mers, counts [C, 2^20+2] or None)."""
        hist = np.zeros(NBUCKETS + 1, dtype=np.uint64)
        reads = np.zeros(max(self.C, 1), dtype=np.uint64)
        tot = ct.c_uint64(0)
        counts = np.zeros((self.C, NBUCKETS + 2), dtype=np.uint32) if want_counts else None
        self._check(lib().mp_count_loaded(self._h, counts.ctypes.data if want_counts else None, _p(hist), _p(reads),
                                          ct.byref(tot)), "mp_count_loaded")
        return hist, reads[: self.C], int(tot.value), counts

    def scan(self, s: int, lo: int, hi: int, seg_bounds, expected: int = 0) -> ScanStats:
        seg = _u64(seg_bounds)
        st = ScanStats()
        self._check(lib().mp_scan(self._h, s, lo, hi, _p(seg), len(seg) - 1, expected, ct.byref(st)),
                    "mp_scan(%d)" % s)
        return st

    def finish(self):
        W = np.zeros(NPART + 1, dtype=np.uint64)
        st = FinishStats()
        self._check(lib().mp_finish(self._h, _p(W), ct.byref(st)), "mp_finish")
        return W, st

    def parent(self) -> np.ndarray:
        out = np.zeros(self.N, dtype=np.uint64)
        self._check(lib().mp_get_parent(self._h, _p(out)), "mp_get_parent")
        return out

    def bad(self) -> np.ndarray:
        out = np.zeros(self.N, dtype=np.uint8)
        self._check(lib().mp_get_bad(self._h, out.ctypes.data), "mp_get_bad")
        return out

    def dump_tuples(self) -> np.ndarray:
        """Last scan's sorted, filtered tuples as an (n, 14) uint8 array (reference layout)."""
        n = ct.c_uint64(0)
        self._check(lib().mp_dump_tuples(self._h, None, 0, ct.byref(n)), "mp_dump_tuples")
        out = np.zeros((int(n.value), 14), dtype=np.uint8)
        self._check(lib().mp_dump_tuples(self._h, out.ctypes.data, n.value, ct.byref(n)), "mp_dump_tuples")
        return out

    def write_partitions(self, out_prefix: Optional[str], T: int, W, il_start=None) -> int:
        Wa = _u64(W)
        il = None
        if il_start is not None:
            ila = np.ascontiguousarray(np.asarray(il_start, dtype=np.int64))
            il = ila.ctypes.data_as(ct.POINTER(ct.c_int64))
        nb = ct.c_uint64(0)
        pre = out_prefix.encode() if out_prefix is not None else None
        self._check(lib().mp_write_partitions(self._h, pre, T, _p(Wa), il, ct.byref(nb)), "mp_write_partitions")
        return int(nb.value)


    # ------------------------------------------------------------ batches of samples
    def batch_load(self, bufs):
        """Upload n independent single-end FASTQ samples (host buffers) back to back."""
        ptrs, sizes = zip(*[_buf_ptr(b) for b in bufs])
        arr = (ct.c_void_p * len(bufs))(*ptrs)
        sz = _u64(sizes)
        self._check(lib().mp_batch_load(self._h, len(bufs), arr, _p(sz)), "mp_batch_load")
        self.C = len(bufs)
        self.N = int(self.first_ids()[-1])

    def batch_load_device(self, dev_ptr: int, nbytes: int, sample_ofs):
        """Samples already back to back in device memory; sample_ofs has n+1 byte offsets."""
        so = _u64(sample_ofs)
        self._check(lib().mp_batch_load_device(self._h, ct.c_void_p(dev_ptr), nbytes, len(so) - 1, _p(so)),
                    "mp_batch_load_device")
        self.C = len(so) - 1
        self.N = int(self.first_ids()[-1])

    def batch_partition(self):
        """(per-sample result dicts, ScanStats, FinishStats) for the loaded batch."""
        res = (SampleResult * self.C)()
        st, fs = ScanStats(), FinishStats()
        self._check(lib().mp_batch_partition(self._h, res, ct.byref(st), ct.byref(fs)), "mp_batch_partition")
        return [r.as_dict() for r in res], st, fs

    def batch_parent(self, s: int) -> np.ndarray:
        ids = self.first_ids()
        out = np.zeros(int(ids[s + 1] - ids[s]), dtype=np.uint32)
        self._check(lib().mp_batch_get_parent(self._h, s, out.ctypes.data), "mp_batch_get_parent")
        return out

    def batch_write(self, out_dir: Optional[str]) -> int:
        nb = ct.c_uint64(0)
        self._check(lib().mp_batch_write(self._h, out_dir.encode() if out_dir is not None else None, ct.byref(nb)),
                    "mp_batch_write")
        return int(nb.value)


class Server:
    """The served form (mp_server): submit jobs at any time, wait for each."""

    def __init__(self, k: int, engines: int = 3, max_jobs_per_batch: int = 64, hbm_budget: int = 0, device: int = 0,
                 min3: bool = False):
        cfg = ServerCfg(device=device, k=k, engines=engines, max_jobs_per_batch=max_jobs_per_batch,
                        hbm_budget=hbm_budget, min3=1 if min3 else 0)
        h = ct.c_void_p()
        rc = lib().mp_server_create(ct.byref(cfg), ct.byref(h))
        if rc:
            raise MPError(rc, "mp_server_create")
        self._h = h
        self._keep = {}  # job -> (buffer, reads hint): the bytes stay alive until waited for

    def submit(self, buf, want_parent: bool = False) -> int:
        ptr, n = _buf_ptr(buf)
        job = ct.c_uint64(0)
        rc = lib().mp_server_submit(self._h, ct.c_void_p(ptr), n, 1 if want_parent else 0, ct.byref(job))
        if rc:
            raise MPError(rc, "mp_server_submit")
        self._keep[job.value] = (buf, n)
        return job.value

    def wait(self, job: int, want_parent: bool = False):
        """(result dict, parent labels or None).  A failed job raises MPError."""
        res = JobResult()
        _, n = self._keep.get(job, (None, 0))
        par = np.zeros(n // 6 + 1, dtype=np.uint32) if want_parent else None  # a record is >= 6 bytes
        rc = lib().mp_server_wait(self._h, job, ct.byref(res), par.ctypes.data if par is not None else None,
                                  par.size if par is not None else 0)
        self._keep.pop(job, None)
        if rc:
            raise MPError(rc, "mp_server_wait(%d)" % job)
        d = res.as_dict()
        if d["status"]:
            raise MPError(d["status"], "job %d failed" % job)
        return d, (par[: d["reads"]].copy() if par is not None else None)

    def clock(self) -> float:
        return float(lib().mp_server_clock(self._h))

    def close(self):
        if self._h:
            lib().mp_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class RunResult:
    W: np.ndarray
    parent: Optional[np.ndarray]
    scans: list
    finish: dict
    hist: np.ndarray
    out_bytes: int


def partition(ctx: Context, S: int = 1, T: int = 1, out_prefix: Optional[str] = None, want_parent: bool = False,
              il_start=None, build_output: bool = True, hist: Optional[np.ndarray] = None) -> RunResult:
    """The metapart stages on reads already loaded in ctx: histogram (unless
    given), plan, S scans, finish, writer (device streams; files too when
    out_prefix is given)."""
    if hist is None:
        hist, reads, total, _ = ctx.count()
    scan, thread = plan_ranges(hist, S, T)
    scans = []
    for s in range(S):
        lo, hi = int(scan[s]), int(scan[s + 1])
        expected = int(hist[hi] - hist[lo]) if hi <= NBUCKETS else 0
        scans.append(ctx.scan(s, lo, hi, thread[s * T:(s + 1) * T + 1], expected).as_dict())
    W, fs = ctx.finish()
    nb = ctx.write_partitions(out_prefix, T, W, il_start) if build_output else 0
    return RunResult(W=W, parent=ctx.parent() if want_parent else None, scans=scans, finish=fs.as_dict(), hist=hist,
                     out_bytes=nb)


# ---------------------------------------------------------------- synthetic data
class SynthCfg(ct.Structure):
    _fields_ = [("seed", ct.c_uint64), ("n_genomes", ct.c_int), ("genome_len_min", ct.c_uint64),
                ("genome_len_max", ct.c_uint64), ("gc_min", ct.c_double), ("gc_max", ct.c_double),
                ("abund_sigma", ct.c_double), ("n_repeat_families", ct.c_int), ("repeat_len", ct.c_int),
                ("repeat_copies", ct.c_int), ("n_reads", ct.c_uint64), ("read_len_min", ct.c_int),
                ("read_len_max", ct.c_int), ("err_rate", ct.c_double), ("interleaved", ct.c_int),
                ("insert_len", ct.c_int), ("n_rate", ct.c_double)]


def synth_lib():
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_PATH):
            raise RuntimeError("libmpsynth.so is not built: run python -m metapart_amd.build")
        L = ct.CDLL(SYNTH_PATH)
        L.mp_synth_preset.argtypes = [ct.c_int, ct.POINTER(SynthCfg)]
        L.mp_synth_generate.argtypes = [ct.POINTER(SynthCfg), ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_size_t)]
        L.mp_synth_free.argtypes = [ct.c_void_p]
        _synth = L
    return _synth


class SynthBuffer:
    """FASTQ bytes held by the C generator (no Python copy)."""

    def __init__(self, ptr: int, nbytes: int):
        self.ptr = ptr
        self.nbytes = nbytes

    def view(self) -> np.ndarray:
        return np.ctypeslib.as_array(ct.cast(self.ptr, ct.POINTER(ct.c_uint8)), shape=(self.nbytes,))

    def tobytes(self) -> bytes:
        return ct.string_at(self.ptr, self.nbytes)

    def free(self):
        if self.ptr:
            synth_lib().mp_synth_free(ct.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synth_cfg(preset: int = 0, **over) -> SynthCfg:
    c = SynthCfg()
    if synth_lib().mp_synth_preset(preset, ct.byref(c)):
        raise ValueError("unknown preset %d" % preset)
    for k, v in over.items():
        if k == "read_len":
            c.read_len_min = c.read_len_max = int(v)
        else:
            setattr(c, k, v)
    return c


def mix_params(n_samples: int = 64, seed: int = 4, reads_min: int = 10_000, reads_max: int = 5_000_000,
               genomes_min: int = 20, genomes_max: int = 300, genome_len_min: int = 100_000,
               genome_len_max: int = 2_000_000, len_min: int = 100, len_max: int = 150, err_min: float = 0.005,
               err_max: float = 0.02, seed0: int = 100):
    """Per-sample generator settings of the 64-sample mix: sample i has its own
    genome pool seeded seed0 + i (genomes_min..genomes_max genomes of
    genome_len_min..genome_len_max bases, GC fraction U(0.3, 0.7) per genome),
    a log-uniform read count (even), one read length in [len_min, len_max] and
    a uniform error rate in [err_min, err_max] (these draws from `seed`)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_samples):
        n = int(np.exp(rng.uniform(np.log(reads_min), np.log(reads_max))))
        n += n & 1
        out.append(dict(seed=seed0 + i, n_genomes=int(rng.integers(genomes_min, genomes_max + 1)),
                        genome_len_min=genome_len_min, genome_len_max=genome_len_max, gc_min=0.3, gc_max=0.7,
                        abund_sigma=1.0, n_reads=n, read_len=int(rng.integers(len_min, len_max + 1)),
                        err_rate=float(rng.uniform(err_min, err_max))))
    return out


def job_params(n_jobs: int = 200, seed: int = 7, reads_min: int = 50_000, reads_max: int = 2_000_000,
               genomes_min: int = 20, genomes_max: int = 100, genome_len_min: int = 500_000,
               genome_len_max: int = 3_000_000, seed0: int = 5000):
    """Partition jobs of the served workload: job j has a fresh genome pool
    seeded seed0 + j (genomes_min..genomes_max genomes of genome_len_min..
    genome_len_max bases), a log-uniform read count (even), 150 bp reads and
    1 % errors (the counts drawn from `seed`)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_jobs):
        n = int(np.exp(rng.uniform(np.log(reads_min), np.log(reads_max))))
        n += n & 1
        out.append(dict(seed=seed0 + i, n_genomes=int(rng.integers(genomes_min, genomes_max + 1)),
                        genome_len_min=genome_len_min, genome_len_max=genome_len_max, abund_sigma=1.0,
                        n_reads=n, read_len=150, err_rate=0.01))
    return out


def synth_many(params, threads: int = 4):
    """synth(0, **p) for every p, a few at a time on host threads (each
    generator call runs its own OpenMP loops and releases the GIL)."""
    import concurrent.futures as cf
    with cf.ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(lambda p: synth(0, **p), params))


def synth(preset: int = 0, **over) -> SynthBuffer:
    c = synth_cfg(preset, **over)
    p = ct.c_void_p()
    n = ct.c_size_t(0)
    rc = synth_lib().mp_synth_generate(ct.byref(c), ct.byref(p), ct.byref(n))
    if rc:
        raise RuntimeError("mp_synth_generate failed (%d)" % rc)
    return SynthBuffer(p.value, n.value)
