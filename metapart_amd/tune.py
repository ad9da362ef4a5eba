"""Autotuning for the device path.

* Pass size: the reference splits the m-mer buckets into numScans passes so
  that one pass's tuples fit memory (metapart.c:3720-3791).  scans_for_budget()
  picks the fewest passes whose largest pass, together with everything that
  stays resident (FASTQ bytes, per-read state, packed streams, the writer's
  buffers), fits an HBM budget.
* Sort tile: tune_sort_tile() times the radix pass with each tile shape the
  library offers on a sample of the workload and returns the fastest.
* Digit layout: tune_sort_digits() times the scans' sort with byte digits and
  with the k-mer's top bits split over the last two passes (k = 29..31; at
  k = 31 7 + 7 bits instead of 8 + 6) and returns the faster for
  mp_cfg.sort_digit_bits.  Digits are at most 8 bits: a 256-bin histogram per
  wave is what fits LDS next to a 6144-tuple staging tile at two tiles per CU.
"""
from __future__ import annotations

import numpy as np

from . import NBUCKETS, Context, MPError, partition, plan_ranges, synth

TUPLE_BYTES = 13           # k-mer u64 + label u32 + adj u8
PER_TUPLE_DEVICE = 31      # as the library plans a scan: ping-pong buffers, bitmaps, status, digit bytes, slack
PER_READ_DEVICE = 74       # record table, labels, flags, sizes, a scan's label snapshot
PER_READ_LATER = 128       # pair-step arrays and writer items, as the library keeps room for them
TILE_SHAPES = (0, 8192, 4096, 6144, 3072)  # 0: 1024 lanes x 6 tuples


def resident_bytes(reads: int, fq_bytes: int, bases: int, kmers: int) -> int:
    """Device bytes that do not depend on the pass size (the library's own
    reckoning when it decides whether a scan fits, mp_api.hip)."""
    return (fq_bytes                      # FASTQ
            + PER_READ_DEVICE * reads     # per-read state
            + (3 * bases) // 8            # 2-bit + N streams
            + bases // 16                 # look-back status of the generating pass
            + (PER_READ_LATER * reads + fq_bytes) // 16 * 17  # pair step and writer (items, output streams)
            + (512 << 20))                # histogram tables, counts per chunk, small arrays


def scans_for_budget(hist: np.ndarray, reads: int, fq_bytes: int, bases: int, budget: int, T: int = 1,
                     max_scans: int = 256):
    """(S, bytes needed) for the fewest passes that fit `budget` device bytes.
    hist is the exclusive prefix over the 2^20 buckets (mp_count_loaded)."""
    hist = np.asarray(hist, dtype=np.uint64)
    kmers = int(hist[NBUCKETS])
    base = resident_bytes(reads, fq_bytes, bases, kmers)
    for S in range(1, max_scans + 1):
        try:
            scan, _ = plan_ranges(hist, S, T)
        except MPError:
            continue
        hi = np.minimum(scan[1:], NBUCKETS).astype(np.int64)
        lo = np.minimum(scan[:-1], NBUCKETS).astype(np.int64)
        biggest = int((hist[hi] - hist[lo]).max())
        # the histogram's low-bit staging (2 B per k-mer) later becomes the second key buffer
        need = base + biggest * PER_TUPLE_DEVICE + max(0, 2 * kmers - 8 * biggest)
        if need <= budget:
            return S, need
    raise MPError(-2, "no pass count up to %d fits %d bytes" % (max_scans, budget))


def tune_sort_tile(k: int = 31, reads: int = 2_000_000, seed: int = 2, device: int = 0) -> dict:
    """Times one partition of a sample workload per tile shape; returns
    {"best": shape, "ms_per_pass": {shape: ms}}."""
    buf = synth(2, n_reads=reads, seed=seed)
    times = {}
    for shape in TILE_SHAPES:
        with Context(k=k, device=device, sort_tile=shape) as ctx:
            best = None
            for _ in range(2):  # the first run allocates
                ctx.load(buf, [0, buf.nbytes])
                r = partition(ctx, S=1, T=1, build_output=False)
                ms = float(r.scans[0]["ms_sort_pass_avg"])
                best = ms if best is None else min(best, ms)
            times[shape] = best
    buf.free()
    return {"best": min(times, key=times.get), "ms_per_pass": times}


def tune_sort_digits(k: int = 31, reads: int = 2_000_000, seed: int = 2, device: int = 0) -> dict:
    """Times the scans' sort (all passes and the local pass) of a sample
    workload with byte digits (8) and with the split top (7); returns
    {"best": sort_digit_bits, "ms_sort": {bits: ms}}."""
    buf = synth(2, n_reads=reads, seed=seed)
    times = {}
    for bits in (8, 7):
        with Context(k=k, device=device, sort_digit_bits=bits) as ctx:
            best = None
            for _ in range(2):  # the first run allocates
                ctx.load(buf, [0, buf.nbytes])
                r = partition(ctx, S=1, T=1, build_output=False)
                ms = float(r.scans[0]["ms_sort"])
                best = ms if best is None else min(best, ms)
            times[bits] = best
    buf.free()
    return {"best": min(times, key=times.get), "ms_sort": times}
