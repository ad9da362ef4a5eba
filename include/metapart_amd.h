/* metapart_amd: MI355X (gfx950) implementation of MetaPrep's read-partitioning
 * path behind a plain C ABI.
 *
 * The reference (vasupsu/MetaPart) has no library API: its boundary is three
 * command-line tools, their files, and link-time "plugins" (getCanonKmers /
 * init / getMmerPrefix).  Each entry point below names the reference code it
 * replaces.  The host tools in metapart_amd/tools keep the reference's
 * command lines and file layouts and call these functions.
 *
 * Conventions
 *  - every call returns 0 on success and a negative MP_E* code on error;
 *    nothing aborts.  mp_last_error() gives the message.
 *  - host buffers are owned by the caller; device memory is owned by the ctx.
 *  - a ctx is used by one host thread at a time; several ctxs may share a GPU.
 *  - calls are stream-ordered on the ctx's stream and synchronous at return.
 *  - read ids and labels are 32-bit: at most 4,294,967,294 reads per run;
 *    the writer (mp_write_partitions, mp_batch_write) takes at most
 *    2,147,483,647 reads and returns MP_EINVAL beyond that.
 */
#ifndef METAPART_AMD_H
#define METAPART_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MP_NBUCKETS 1048576u /* 4^10 m-mer buckets, m = 10 (genHistograms.c:105) */
#define MP_NPART 16          /* output partitions (metapart.c:4016) */
#define MP_PAIR_GATE 1000u   /* size gate of the pair step (metapart.c:4024) */
#define MP_MAX_SAMPLES 256   /* samples per batch (one 8-bit sort digit) */

enum {
    MP_OK = 0,
    MP_EINVAL = -1,   /* bad argument / malformed input */
    MP_ENOMEM = -2,   /* host or device allocation failed, or HBM budget exceeded */
    MP_EPLAN = -3,    /* degenerate range plan the reference cannot run either */
    MP_ECOUNT = -4,   /* enumeration disagrees with the histogram sizes */
    MP_EDEVICE = -5,  /* HIP runtime error */
    MP_EIO = -6,      /* file error */
    MP_ESTATE = -7    /* call out of order */
};

/* ---------------------------------------------------------------- host side */

/* Chunk index: writes <files[0]>.idx with num_chunks+1 records of the
 * reference's fileIndex layout (4,194,344 bytes each; unspecified bytes are 0).
 * Replaces genFastqIdx main (genFastqIdx.c:142-376).  nPE counts file PAIRS. */
int mp_index_fastq(int num_chunks, int nPE, int nSE, int nIL, const char *const *files, int nfiles);

/* The same chunk boundaries for one in-memory FASTQ file (single-end, or
 * interleaved pairs when interleaved != 0): chunk_ofs receives up to
 * num_chunks+1 byte offsets (the last is nbytes), *n_chunks the chunk count
 * (smaller than num_chunks when the reference's near-end-of-file rule stops
 * early, genFastqIdx.c:219-225). */
int mp_index_buffer(const char *buf, uint64_t nbytes, int num_chunks, int interleaved, uint64_t *chunk_ofs,
                    int *n_chunks);

/* Range planning with one MPI rank: scan[S+1] and thread[S*T+1] bucket
 * boundaries from the exclusive-prefix histogram hist[2^20+1].
 * Replaces metapart.c:3720-3791.  Returns MP_EPLAN when a walk would write
 * past its array or leaves a thread range empty or out of order. */
int mp_plan_ranges(const uint64_t *hist, int S, int T, uint64_t *scan, uint64_t *thread);

/* ---------------------------------------------------------------- device side */

typedef struct mp_ctx mp_ctx;

typedef struct {
    int device;           /* HIP device ordinal */
    int k;                /* k-mer length, 11..32 */
    int min3;             /* 0: every k-mer; 1: three window minimizers per read */
    uint64_t hbm_budget;  /* device bytes this ctx may hold; 0 = 90 % of free memory */
    int sort_digit_bits;  /* scans' radix digits: 8 = bytes; 0 or 7 = bytes, except that at k = 29..31 the
                             k-mer's bits above bit 48 are split evenly over the last two passes (k = 31:
                             7 + 7 instead of 8 + 6); anything else is MP_EINVAL */
    int sort_tile;        /* tuples per radix tile: 0 (1024 lanes x 6), 3072, 4096, 6144 (512 lanes) or 8192 */
    int sort_graph;       /* 1: a radix pass chain of counted passes only (a batch's) is captured in a HIP graph
                             and replayed while the sort's shape and buffers repeat; 0: launched one by one */
} mp_cfg;

typedef struct {
    uint64_t tuples;          /* tuples enumerated in the scan */
    uint64_t bad_labels;      /* labels marked BAD by the filter in this scan */
    uint64_t zeroed_runs;     /* runs whose k-mer was zeroed (both sides branch) */
    uint64_t evaluated_runs;  /* runs the filter evaluated */
    int quirk_segments;       /* segments whose first run joined tuple 0's label */
    int sort_passes;          /* radix passes executed */
    float ms_enumerate, ms_sort, ms_filter, ms_union; /* device time */
    float ms_sort_pass_avg;   /* mean duration of one radix scatter kernel (passes after the first) */
    float ms_sort_hist;       /* the digit-histogram read before the passes */
    float ms_sort_local;      /* the local pass ranking tuples inside high-bit groups (0: all passes global) */
    uint64_t sort_long_groups;/* groups too long for the local pass, sorted by their own global passes */
    float ms_sort_gen;        /* the first radix pass, generating the tuples from the reads (0: enumerated) */
    int parts;                /* 0: one pass; else the scan did not fit the budget and went twice over this many bucket parts */
    uint64_t sort_local_moved;/* tuples the in-place local pass moved (the others were already at their rank) */
    float ms_sort_count;      /* digit counts and offsets of the passes that take them from a count (not look-back) */
    int sort_graph;           /* the radix pass chain: 0 launched pass by pass, 1 captured in a graph now, 2 replayed */
} mp_scan_stats;

typedef struct {
    uint64_t pair_unions;    /* unions made by the pair step */
    uint64_t pair_candidates;/* pairs that needed the ordered pass */
    uint64_t components;     /* roots after the pair step */
    uint64_t largest;        /* size of the largest component */
    float ms_device, ms_host;
} mp_finish_stats;

int mp_ctx_create(const mp_cfg *cfg, mp_ctx **out);
int mp_ctx_destroy(mp_ctx *ctx);
const char *mp_last_error(const mp_ctx *ctx);
/* Device bytes currently held by the ctx. */
uint64_t mp_ctx_device_bytes(const mp_ctx *ctx);

/* m-mer histograms of C chunks of FASTQ bytes.  Chunk c is
 * fq[chunk_ofs[c] .. chunk_ofs[c+1]) and starts at a record.  counts receives
 * C x (2^20+2) u32 (the kmerFreqCount of each chunk record, last two 0),
 * reads_per_chunk C values; *total_kmers the sum.
 * Replaces genHistograms.c:161-265 with the plugin getCanonKmers of
 * countMmers.c:83-217 (min3 = 0) or countMmers_3minimizers.c:83-271. */
int mp_count_chunks(mp_ctx *ctx, const char *fq, uint64_t nbytes, const uint64_t *chunk_ofs, int C,
                    uint32_t *counts, uint64_t *reads_per_chunk, uint64_t *total_kmers);

/* The same histograms without allocating: no counterpart in the reference;
 * for callers that cannot allocate device memory (request loops, harnesses
 * that forbid hipMalloc).  The FASTQ bytes are already in device memory
 * (16-byte aligned, on `device`); counts, reads_per_chunk and total_kmers are
 * host buffers as in mp_count_chunks, with k and min3 as in mp_cfg.  Every
 * device buffer comes from one fixed array of MP_NOALLOC_SCRATCH_BYTES in
 * device memory, which each GPU the library runs on holds until the process
 * ends; the first call on a device also creates the stream and events the
 * calls use.  No call allocates or frees device or pinned memory.  Inputs
 * above MP_NOALLOC_MAX_FASTQ bytes or MP_NOALLOC_MAX_CHUNKS chunks return
 * MP_EINVAL, and an input whose records need more than the array returns
 * MP_ENOMEM.  Calls share the device's array: calls that overlap on one GPU
 * race (serialise them). */
#define MP_NOALLOC_SCRATCH_BYTES (128ull << 20)
#define MP_NOALLOC_MAX_FASTQ (32ull << 20)
#define MP_NOALLOC_MAX_CHUNKS 8
int mp_count_chunks_noalloc(int device, int k, int min3, const void *d_fq, uint64_t nbytes,
                            const uint64_t *chunk_ofs, int C, uint32_t *counts, uint64_t *reads_per_chunk,
                            uint64_t *total_kmers);

/* Uploads the FASTQ bytes of a run (C chunks as above) and parses every
 * record; chunk_first_id[c] is the read id of chunk c's first record
 * (KmerIdx readId/readIdExtra), chunk_first_id[C] the read count.
 * Reads stay resident across scans.  An odd read count loads and can be
 * counted (as genHistograms counts it); mp_scan refuses it with MP_EINVAL
 * (the reference's pair step reads id N of such a run and crashes).
 * Replaces loadChunk / getNextRead (metapart.c:296-500) and the scan-0
 * parent initialisation (metapart.c:3933-3954). */
int mp_load_reads(mp_ctx *ctx, const char *fq, uint64_t nbytes, const uint64_t *chunk_ofs,
                  const uint64_t *chunk_first_id, int C);

/* Same, for FASTQ bytes already in device memory (not copied; must stay valid
 * until the next mp_load_* or mp_ctx_destroy). */
int mp_load_reads_device(mp_ctx *ctx, const void *d_fq, uint64_t nbytes, const uint64_t *chunk_ofs,
                         const uint64_t *chunk_first_id, int C);

/* m-mer histograms of the reads already loaded (no upload, no re-parse):
 * counts as in mp_count_chunks (may be NULL), hist the exclusive prefix over
 * buckets, 2^20+1 u64 as the kmerHist file holds it (may be NULL),
 * reads_per_chunk C values (may be NULL).
 * Replaces genHistograms.c:161-344 for a run that stays on one device. */
int mp_count_loaded(mp_ctx *ctx, uint32_t *counts, uint64_t *hist, uint64_t *reads_per_chunk, uint64_t *total_kmers);

/* Read id of the first record of every chunk of the loaded run (C+1 values,
 * the last is the read count): what the KmerIdx readId fields hold. */
int mp_chunk_first_ids(mp_ctx *ctx, uint64_t *ids);

/* One scan over buckets [lo, hi) split into T segments at seg_bounds[0..T]
 * (thread ranges of the plan): enumerate, sort, filter, union, compress.
 * Replaces generateKmersInScan + getCanonKmers of enumerateKmers.c:127-371 or
 * enumerateKmers_3minimizers.c:129-474, sortKmerList, markMischiefKmers,
 * unionFindPar2 / unionFind and printStats (metapart.c:1192-2800, 3227-3308).
 * expected_tuples (may be 0 = unchecked) is the histogram's count for the
 * scan; a mismatch returns MP_ECOUNT as the reference would overflow. */
int mp_scan(mp_ctx *ctx, int s, uint64_t lo, uint64_t hi, const uint64_t *seg_bounds, int T,
            uint64_t expected_tuples, mp_scan_stats *stats);

/* After the last scan: component sizes, the ordered pair step with the size
 * gate, sizes again, and the 16-way partition plan W[0..16].
 * Replaces metapart.c:4010-4056 and unionFindPeReads (2801-2867). */
int mp_finish(mp_ctx *ctx, uint64_t *W, mp_finish_stats *stats);

/* Parent label of every read (read count entries): the component root after
 * the last scan (a BAD read is its own parent here, where the reference
 * stores the sentinel 0xBADBADBAD), or the final root after mp_finish. */
int mp_get_parent(mp_ctx *ctx, uint64_t *parent);

/* BAD flag of every read (1 = marked by the frequency filter and not yet
 * reset by the pair step).  Debug and parity only. */
int mp_get_bad(mp_ctx *ctx, uint8_t *bad);

/* The last scan's sorted, filtered tuple array in the reference's 14-byte
 * layout (enumerateKmers.c:352-365: k-mer u64, adj u8, id/UINT32_MAX u8,
 * id%UINT32_MAX u32).  Debug and parity only. */
int mp_dump_tuples(mp_ctx *ctx, uint8_t *dst, uint64_t cap, uint64_t *n);

/* Partitioned FASTQ writer.  Builds the 16*T output streams on the device
 * (record bytes copied verbatim, pairs duplicated when their roots differ)
 * and writes OUT/Rank0/FromTid<t>_ToRank<p>.paired.fastq.
 * il_start[c] < 0: chunk c has no interleaved part; otherwise the byte offset
 * (within the chunk) from which records pair up, as loadChunk leaves it.
 * out_prefix NULL: build the streams on the device only (timing).
 * Replaces writeOutput3 (metapart.c:956-1066). */
int mp_write_partitions(mp_ctx *ctx, const char *out_prefix, int T, const uint64_t *W, const int64_t *il_start,
                        uint64_t *bytes_out);

/* ---------------------------------------------------------------- batches
 * Batched form.  The reference partitions one sample per process (metapart.c
 * main); these entry points have no counterpart there and serve callers with
 * many independent samples.  n samples, each a complete single-end FASTQ with
 * an even, non-zero record count, go through one launch sequence: one parse,
 * one enumeration, one sort whose last pass groups the tuples by sample, one
 * filter/union pass with one segment per sample, one pair step, and per-sample
 * plans.  Every sample's labels, plan and output files are those of a plain
 * run of that sample alone with one chunk, one scan and one segment
 * (numThreads = 1): mp_load_reads with C = 1, mp_scan over [0, 2^20),
 * mp_finish, mp_write_partitions with T = 1. */
typedef struct {
    uint64_t reads;               /* records of the sample */
    uint64_t tuples;              /* k-mer tuples of the sample */
    uint64_t components;          /* components after the pair step */
    uint64_t largest;             /* largest component */
    uint64_t W[MP_NPART + 1];     /* partition plan over the sample's own read ids */
} mp_sample_result;

/* Loads n (<= MP_MAX_SAMPLES) host FASTQ buffers back to back into the ctx. */
int mp_batch_load(mp_ctx *ctx, int n, const char *const *fq, const uint64_t *nbytes);

/* Same for samples already back to back in device memory: sample i holds
 * bytes [sample_ofs[i], sample_ofs[i+1]) of d_fq (sample_ofs[n] = nbytes). */
int mp_batch_load_device(mp_ctx *ctx, const void *d_fq, uint64_t nbytes, int n, const uint64_t *sample_ofs);

/* Partitions every loaded sample; res (n entries) may be NULL. */
int mp_batch_partition(mp_ctx *ctx, mp_sample_result *res, mp_scan_stats *scan_stats, mp_finish_stats *finish_stats);

/* Final labels of sample s over its own read ids (reads entries). */
int mp_batch_get_parent(mp_ctx *ctx, int s, uint32_t *parent);

/* Output files OUT/sample<s>/Rank0/FromTid0_ToRank<p>.paired.fastq of every
 * sample (out_dir NULL: build the streams on the device only). */
int mp_batch_write(mp_ctx *ctx, const char *out_dir, uint64_t *bytes_out);

/* ---------------------------------------------------------------- served form
 * A request loop inside the library (no counterpart in the reference, which
 * runs one sample per process; these entry points serve callers that submit
 * partition jobs over time).  Jobs queue as they arrive; `engines` contexts,
 * each with its own stream and an equal share of hbm_budget, run on host
 * threads.  An idle engine takes, in arrival order, every queued job whose
 * estimated device footprint fits its share (at most max_jobs_per_batch) and
 * partitions them as one batch, so jobs arriving while others run are admitted
 * as engines and memory free up, and batches overlap on the device.  A job's
 * result equals a plain run of its sample alone (see the batched form). */
typedef struct mp_server mp_server;

typedef struct {
    int device;
    int k;
    int engines;              /* concurrent contexts (0: 3) */
    int max_jobs_per_batch;   /* 0: 64; at most MP_MAX_SAMPLES */
    uint64_t hbm_budget;      /* device bytes for all engines (0: 3/4 of free memory) */
    int min3;                 /* as mp_cfg: 0 every k-mer, 1 three window minimizers per read */
} mp_server_cfg;

typedef struct {
    uint64_t job;
    int status;               /* 0 or the MP_E* code of the failed job */
    int batch_jobs;           /* jobs in the batch that ran it */
    uint64_t reads, tuples, components, largest;
    uint64_t W[MP_NPART + 1];
    double t_submit, t_start, t_done;  /* seconds on the server clock (mp_server_clock) */
} mp_job_result;

int mp_server_create(const mp_server_cfg *cfg, mp_server **out);

/* Queues one single-end FASTQ sample (even, non-zero record count).  fq must
 * stay valid until mp_server_wait returns for the job. */
int mp_server_submit(mp_server *srv, const char *fq, uint64_t nbytes, int want_parent, uint64_t *job);

/* Blocks until the job is done; parent (cap entries, may be NULL) receives
 * its labels over its own read ids if it was submitted with want_parent.
 * A job is waited for once: a second wait on the same job, from any thread,
 * returns MP_ESTATE while the first is pending and MP_EINVAL after it. */
int mp_server_wait(mp_server *srv, uint64_t job, mp_job_result *res, uint32_t *parent, uint64_t cap);

double mp_server_clock(const mp_server *srv);

/* Runs every queued job to completion, then frees the engines. */
int mp_server_destroy(mp_server *srv);

#ifdef __cplusplus
}
#endif
#endif
