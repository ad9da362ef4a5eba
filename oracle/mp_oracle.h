/* CPU oracle: a plain-C restatement of MetaPrep's read-partitioning path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * (or the timed CPU baseline).  The product (metapart_amd) never calls it.
 *
 * Every function restates the reference loop it cites (file:line of
 * vasupsu/MetaPart) single-threaded, with the OpenMP thread count T of the
 * reference turned into T "virtual segments" so that every T-dependent
 * behaviour (filter segments, first-run union quirk, output file split) is
 * reproduced exactly.  MPI is fixed at one rank.
 *
 * Pinning: each stage is checked by tests/test_oracle_*.py against the real
 * reference built from its sources (oracle/_ref, golden fixtures under
 * tests/golden/), against closed forms (canonical k-mers, prefix sums,
 * connected components by brute force) and invariants.
 */
#ifndef MP_ORACLE_H
#define MP_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NBUCKETS 1048576u
#define ORC_BAD 0xBADBADBADull

/* On-disk chunk record: same byte layout as the reference's fileIndex
 * (genFastqIdx.c:11-18): 4,194,344 bytes on x86-64. */
typedef struct {
    uint64_t offset;
    uint64_t offset2;
    uint8_t  readIdExtra;
    uint8_t  pad0[3];
    uint32_t readId;
    int32_t  fileNo;
    uint32_t kmerFreqCount[ORC_NBUCKETS + 2];
    uint8_t  pad1[4];
} orc_fileindex;

/* genFastqIdx: writes <files[0]>.idx (genFastqIdx.c:142-376). nPE = number of
 * file PAIRS (argv[2]); files are in argv order.  Returns 0, <0 on error. */
int orc_gen_fastq_idx(int num_chunks, int nPE, int nSE, int nIL, const char *const *files, int nfiles);

/* genHistograms_metapart[min]: reads <files[0]>.idx, writes KmerIdx and kmerHist
 * (genHistograms.c:82-344).  min3 selects the 3-minimizer counter. */
int orc_gen_histograms(int K, int nPE, int nSE, int nIL, const char *const *files, int nfiles, int min3);

/* metapart[min] -o out bufSz K S T nPE nSE nIL files... (metapart.c:3544-4150).
 * Writes out/Rank0/FromTid<t>_ToRank<p>.paired.fastq.  If parent_out is not
 * NULL it receives the final parent label of every read (numReads entries,
 * allocated by the oracle, free with orc_free). */
int orc_metapart(const char *out_prefix, uint64_t buf_size, int K, int S, int T, int nPE, int nSE, int nIL,
                 const char *const *files, int nfiles, int min3, uint64_t **parent_out, uint64_t *num_reads_out);

/* ---- stage-level entry points used by the unit tests ------------------- */

/* Canonical k-mers of one read, reference counter semantics
 * (countMmers.c:83-217 / countMmers_3minimizers.c:83-271): adds to counts
 * (2^20+2 entries; may be NULL) and returns the number of k-mers counted. */
int orc_count_read(const char *seq, int len, int K, int min3, uint32_t *counts);

/* Tuples of one read (enumerateKmers.c:127-371 / _3minimizers.c:129-474) for
 * buckets in [lo,hi): writes up to cap 14-byte tuples into out, returns count. */
int orc_enum_read(const char *seq, int len, int K, int min3, uint64_t lo, uint64_t hi, uint64_t label,
                  uint8_t *out, int cap);

/* Range planning (metapart.c:3720-3791) with P = 1: scan[S+1], thread[S*T+1]. */
int orc_plan_ranges(const uint64_t *hist, int S, int T, uint64_t *scan, uint64_t *thread);

/* Debug dump of the last metapart run: the sorted, filtered tuple array of scan s
 * (14-byte reference layout), set by orc_set_dump_scan before orc_metapart. */
void orc_set_dump_scan(int scan);
int orc_get_dump(uint8_t **tuples, uint64_t *n, uint64_t **parent_after, uint64_t *nreads);
/* Runs the filter of scan `scan` of the last orc_metapart evaluated (one per
 * k-mer change inside a segment) and zeroed (both neighbour sides branch). */
int orc_get_filter_counts(int scan, uint64_t *evaluated, uint64_t *zeroed);

/* last error message */
const char *orc_last_error(void);
void orc_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
