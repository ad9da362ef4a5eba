"""Shared test helpers: golden fixtures, tool-chain runners and oracle bindings.

The oracle (oracle/liboracle.so, oracle/mp_oracle) is the CPU checker; the
GPU path is exercised only through the C ABI (libmetapart_amd.so) and the
drop-in tools built on it.
"""
import ctypes as ct
import hashlib
import json
import os
import shutil
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")
BIN = os.path.join(ROOT, "metapart_amd", "bin")
ORACLE_CLI = os.path.join(ROOT, "oracle", "mp_oracle")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REC = 4194344
U32MAX = 4294967295
BAD = 0xBADBADBAD


def golden():
    with open(os.path.join(GOLDEN_DIR, "golden.json")) as f:
        return json.load(f)


def golden_case(name):
    for c in golden()["cases"]:
        if c["name"] == name:
            return c
    raise KeyError(name)


def make_input(name, dst):
    """Writes the input a golden case names (reference fixture or seeded synthetic FASTQ)."""
    if name.endswith(".fastq"):
        shutil.copy(os.path.join(GOLDEN_DIR, name), dst)
        os.chmod(dst, 0o644)
    else:
        args = list(golden()["inputs"][name])
        keep = None
        if "--keep-records" in args:  # the generator writes even counts: cut its file
            i = args.index("--keep-records")
            keep = int(args[i + 1])
            del args[i:i + 2]
        subprocess.run([os.path.join(BIN, "mp_synth")] + args + ["-o", dst], check=True)
        if keep is not None:
            lines = open(dst, "rb").read().split(b"\n")
            open(dst, "wb").write(b"\n".join(lines[:4 * keep]) + b"\n")


def masked_sha(path):
    b = bytearray(open(path, "rb").read())
    for r in range(len(b) // REC):
        o = r * REC
        if r == 0:
            b[o + 16] = 0
        b[o + 17:o + 20] = b"\0\0\0"
        b[o + REC - 4:o + REC] = b"\0\0\0\0"
    return hashlib.sha256(bytes(b)).hexdigest()


def sha(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for blk in iter(lambda: f.read(64 << 20), b""):
            h.update(blk)
    return h.hexdigest()


def hash_outputs(d):
    out = {}
    for f in sorted(os.listdir(d)):
        if f.startswith("in.fastq."):
            p = os.path.join(d, f)
            out[f[len("in.fastq."):]] = masked_sha(p) if "idx" in f.lower() else sha(p)
    od = os.path.join(d, "out", "Rank0")
    for f in sorted(os.listdir(od)):
        out["out/" + f] = sha(os.path.join(od, f))
    return out


def golden_large():
    with open(os.path.join(GOLDEN_DIR, "golden_large.json")) as f:
        return json.load(f)


def make_large_input(name, dst):
    """Writes the input of a workload-size golden case (seeded generator)."""
    args = golden_large()["inputs"][name]
    subprocess.run([os.path.join(BIN, "mp_synth")] + args + ["-o", dst], check=True)


def run_chain(kind, case, d, timeout=600, input_path=None, cli=ORACLE_CLI):
    """Runs genFastqIdx, genHistograms_metapart[min], metapart[min] of `kind`
    ('oracle' = CPU restatement, or the oracle build at `cli`; 'gpu' = this
    build's tools) in directory d, on the case's input (or on input_path,
    hard-linked into d)."""
    os.makedirs(os.path.join(d, "out", "Rank0"), exist_ok=True)
    if input_path:
        os.link(input_path, os.path.join(d, "in.fastq"))
    else:
        make_input(case["input"], os.path.join(d, "in.fastq"))
    a = ["0", "0", "1"] if case["mode"] == "il" else ["0", "1", "0"]
    v = case["variant"]
    C, K, S, T = str(case["C"]), str(case["K"]), str(case["S"]), str(case["T"])
    if kind == "oracle":
        cmds = [[cli, "genFastqIdx", C] + a + ["in.fastq"],
                [cli, "genHistograms" + v, K] + a + ["in.fastq"],
                [cli, "metapart" + v, "-o", "./out", "2000000000", K, S, T] + a + ["in.fastq"]]
    else:
        cmds = [[os.path.join(BIN, "genFastqIdx"), C] + a + ["in.fastq"],
                [os.path.join(BIN, "genHistograms_metapart" + v), K] + a + ["in.fastq"],
                [os.path.join(BIN, "metapart" + v), "-o", "./out", "2000000000", K, S, T] + a + ["in.fastq"]]
    for c in cmds:
        r = subprocess.run(c, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError("%s failed (%d):\n%s" % (" ".join(c), r.returncode, r.stdout[-3000:]))
    return hash_outputs(d)


def run_degenerate(kind, case, d, timeout=300):
    """An input whose metapart step the reference cannot run (golden.json
    "degenerate"): genFastqIdx and genHistograms of `kind` must succeed; returns
    the hashes of the in.fastq.* files they wrote and metapart's completed
    process (not checked here)."""
    os.makedirs(os.path.join(d, "out", "Rank0"), exist_ok=True)
    make_input(case["input"], os.path.join(d, "in.fastq"))
    a = ["0", "0", "1"] if case["mode"] == "il" else ["0", "1", "0"]
    v, C, K = case["variant"], str(case["C"]), str(case["K"])
    if kind == "oracle":
        pre = [[ORACLE_CLI, "genFastqIdx", C] + a + ["in.fastq"], [ORACLE_CLI, "genHistograms" + v, K] + a + ["in.fastq"]]
        last = [ORACLE_CLI, "metapart" + v, "-o", "./out", "2000000000", K, "1", "1"] + a + ["in.fastq"]
    else:
        pre = [[os.path.join(BIN, "genFastqIdx"), C] + a + ["in.fastq"],
               [os.path.join(BIN, "genHistograms_metapart" + v), K] + a + ["in.fastq"]]
        last = [os.path.join(BIN, "metapart" + v), "-o", "./out", "2000000000", K, "1", "1"] + a + ["in.fastq"]
    for c in pre:
        subprocess.run(c, cwd=d, check=True, stdout=subprocess.DEVNULL, timeout=timeout)
    h = {}
    for f in sorted(os.listdir(d)):
        if f.startswith("in.fastq."):
            p = os.path.join(d, f)
            h[f[len("in.fastq."):]] = masked_sha(p) if "idx" in f.lower() else sha(p)
    r = subprocess.run(last, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=timeout)
    return h, r


def diff_hashes(got, want):
    bad = sorted(k for k in set(got) | set(want) if got.get(k) != want.get(k))
    return bad


# ---------------------------------------------------------------- oracle library
_orc = None


def oracle():
    global _orc
    if _orc is None:
        L = ct.CDLL(ORACLE_LIB)
        cpp = ct.POINTER(ct.c_char_p)
        L.orc_gen_fastq_idx.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, cpp, ct.c_int]
        L.orc_gen_histograms.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, cpp, ct.c_int, ct.c_int]
        L.orc_metapart.argtypes = [ct.c_char_p, ct.c_uint64, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                   ct.c_int, cpp, ct.c_int, ct.c_int, ct.POINTER(ct.POINTER(ct.c_uint64)),
                                   ct.POINTER(ct.c_uint64)]
        L.orc_count_read.argtypes = [ct.c_char_p, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p]
        L.orc_enum_read.argtypes = [ct.c_char_p, ct.c_int, ct.c_int, ct.c_int, ct.c_uint64, ct.c_uint64, ct.c_uint64,
                                    ct.c_void_p, ct.c_int]
        L.orc_plan_ranges.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p]
        L.orc_set_dump_scan.argtypes = [ct.c_int]
        L.orc_get_filter_counts.argtypes = [ct.c_int, ct.POINTER(ct.c_uint64), ct.POINTER(ct.c_uint64)]
        L.orc_get_dump.argtypes = [ct.POINTER(ct.POINTER(ct.c_uint8)), ct.POINTER(ct.c_uint64),
                                   ct.POINTER(ct.POINTER(ct.c_uint64)), ct.POINTER(ct.c_uint64)]
        L.orc_last_error.restype = ct.c_char_p
        L.orc_free.argtypes = [ct.c_void_p]
        _orc = L
    return _orc


def _argv(files):
    arr = (ct.c_char_p * len(files))(*[f.encode() for f in files])
    return ct.cast(arr, ct.POINTER(ct.c_char_p))


def oracle_run(fq_path, C, K, S, T, interleaved=False, min3=False, dump_scan=-1, out_dir=None):
    """Oracle pipeline on one file.  Returns dict with final parent (N), the
    filter's evaluated and zeroed run counts per scan, and, if dump_scan >= 0,
    the sorted filtered tuples (n, 14) and the parent after that scan (BAD
    reads as 0xBADBADBAD)."""
    L = oracle()
    a = (0, 0, 1) if interleaved else (0, 1, 0)
    files = _argv([fq_path])
    rc = L.orc_gen_fastq_idx(C, a[0], a[1], a[2], files, 1)
    assert rc == 0, L.orc_last_error()
    rc = L.orc_gen_histograms(K, a[0], a[1], a[2], files, 1, 1 if min3 else 0)
    assert rc == 0, L.orc_last_error()
    d = out_dir or os.path.dirname(fq_path)
    os.makedirs(os.path.join(d, "out", "Rank0"), exist_ok=True)
    L.orc_set_dump_scan(dump_scan)
    par = ct.POINTER(ct.c_uint64)()
    nr = ct.c_uint64(0)
    rc = L.orc_metapart(os.path.join(d, "out").encode(), 2000000000, K, S, T, a[0], a[1], a[2], files, 1,
                        1 if min3 else 0, ct.byref(par), ct.byref(nr))
    assert rc == 0, L.orc_last_error()
    res = {"parent": np.ctypeslib.as_array(par, shape=(nr.value,)).copy()}
    L.orc_free(par)
    # per scan: (runs the filter evaluated, runs it zeroed)
    res["filter_counts"] = []
    for sc in range(S):
        ev, zr = ct.c_uint64(0), ct.c_uint64(0)
        assert L.orc_get_filter_counts(sc, ct.byref(ev), ct.byref(zr)) == 0
        res["filter_counts"].append((ev.value, zr.value))
    if dump_scan >= 0:
        tp = ct.POINTER(ct.c_uint8)()
        n = ct.c_uint64(0)
        pp = ct.POINTER(ct.c_uint64)()
        nn = ct.c_uint64(0)
        L.orc_get_dump(ct.byref(tp), ct.byref(n), ct.byref(pp), ct.byref(nn))
        res["tuples"] = np.ctypeslib.as_array(tp, shape=(n.value, 14)).copy() if n.value else np.zeros((0, 14), np.uint8)
        res["parent_scan"] = np.ctypeslib.as_array(pp, shape=(nn.value,)).copy()
        L.orc_free(tp)
        L.orc_free(pp)
        L.orc_set_dump_scan(-1)
    return res


def read_kmeridx(path):
    """(headers, counts[C+1, 2^20+2]) of a KmerIdx file."""
    raw = np.fromfile(path, dtype=np.uint8)
    n = raw.size // REC
    recs = raw.reshape(n, REC)
    counts = recs[:, 28:28 + 4 * (2 ** 20 + 2)].copy().view(np.uint32)
    hdr = [dict(offset=int(r[0:8].view(np.uint64)[0]), readIdExtra=int(r[16]), readId=int(r[20:24].view(np.uint32)[0]),
                fileNo=int(r[24:28].view(np.int32)[0])) for r in recs]
    return hdr, counts


def oracle_sample(job):
    """One sample through the oracle in a process of its own (for process
    pools: nothing here touches the GPU).  job = (fastq path, C, K, S, T,
    out_dir); returns (parent labels as uint32, {file name: sha256})."""
    path, C, K, S, T, out_dir = job
    r = oracle_run(path, C, K, S, T, out_dir=out_dir)
    od = os.path.join(out_dir, "out", "Rank0")
    files = {f: sha(os.path.join(od, f)) for f in sorted(os.listdir(od))}
    return r["parent"].astype(np.uint32), files


def oracle_samples(jobs, workers=8):
    """oracle_sample over many jobs in fresh interpreter processes."""
    import multiprocessing as mpr
    with mpr.get_context("spawn").Pool(processes=max(1, min(workers, len(jobs)))) as pool:
        return pool.map(oracle_sample, jobs)
