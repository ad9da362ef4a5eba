/* CPU oracle: plain-C restatement of the MetaPrep read-partitioning path.
 *
 * TEST INFRASTRUCTURE ONLY (see mp_oracle.h).  The product never links this.
 *
 * Citations are file:line in the reference repository vasupsu/MetaPart.
 * Where the reference has undefined behaviour (writes past arrays, reads of
 * uninitialised stack bytes, empty thread ranges that desynchronise its two
 * offset computations, 3-minimizer buffer overflows at S > 1) the oracle
 * stops with an error instead of guessing; the product does the same, and
 * the tests check that both refuse the same inputs.
 */
#include "mp_oracle.h"

#include <errno.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#define U32MAX 4294967295ull
#define REC_SIZE ((uint64_t)sizeof(orc_fileindex))

_Static_assert(sizeof(orc_fileindex) == 4194344, "fileIndex record must be 4,194,344 bytes");

static char g_err[1024];
static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}
const char *orc_last_error(void) { return g_err; }
void orc_free(void *p) { free(p); }

static int file_size(const char *path, uint64_t *sz)
{
    struct stat st;
    if (stat(path, &st) != 0) return -1;
    *sz = (uint64_t)st.st_size;
    return 0;
}

/* ======================================================================
 * genFastqIdx (genFastqIdx.c)
 * ====================================================================== */

/* genFastqIdx.c:28-53 */
static int get_file_index(int chunk, uint64_t per, int nfiles, const uint64_t *fsize, int numPEFiles,
                          int32_t *fileNo, uint64_t *offset)
{
    int cur = 0;
    uint64_t rem = per * (uint64_t)chunk;
    int factor = cur < numPEFiles ? 2 : 1;
    while (cur < nfiles && rem >= fsize[cur] * (uint64_t)factor) {
        rem -= fsize[cur] * (uint64_t)factor;
        cur += cur < numPEFiles ? 2 : 1;
        factor = cur < numPEFiles ? 2 : 1;
    }
    if (cur >= nfiles) return fail(-1, "chunk %d offset past the last file", chunk);
    *fileNo = cur;
    *offset = rem / (uint64_t)factor;
    return 0;
}

/* genFastqIdx.c:56-139: offset of the mate record in file fileNo+1 */
static int find_ofs_other_end(const char *const *files, const uint64_t *fsize, int fileNo, uint64_t offset,
                              const char *readName, uint64_t *out)
{
    uint64_t delta = fsize[fileNo] * 10 / 1000;
    uint64_t startOfs = offset >= delta ? offset - delta : 0;
    uint64_t len = fsize[fileNo + 1] - startOfs;
    if (offset + delta < fsize[fileNo + 1]) len = offset + delta - startOfs;
    char *b = (char *)calloc(len + 4096, 1);
    if (!b) return fail(-2, "oom");
    FILE *fp = fopen(files[fileNo + 1], "rb");
    if (!fp) { free(b); return fail(-1, "cannot open %s", files[fileNo + 1]); }
    fseek(fp, (long)startOfs, SEEK_SET);
    size_t got = fread(b, 1, len, fp);
    (void)got;
    fclose(fp);
    uint64_t i = 0;
    while (i < len && (b[i] != '\n' || b[i + 1] != '+')) i++;
    if (i >= len) { free(b); return fail(-1, "mate search: no record boundary"); }
    i++;
    while (b[i] != '\n') i++;
    i++;
    while (b[i] != '\n') i++;
    i++;
    if (i >= len) { free(b); return fail(-1, "mate search: window too small"); }
    char cur[256];
    for (;;) {
        int n = 0;
        uint64_t j = i;
        while (b[j] != '\n' && n < 255) cur[n++] = b[j++];
        if (n < 2) { free(b); return fail(-1, "mate search: short name"); }
        cur[n - 2] = '\0';
        if (strcmp(cur, readName) == 0 || i >= len) break;
        for (int l = 0; l < 4; l++) { while (b[i] != '\n') i++; i++; }
        if (i >= len) break;
    }
    free(b);
    *out = startOfs + i;
    return 0;
}

/* genFastqIdx.c:142-376 */
int orc_gen_fastq_idx(int num_chunks, int nPE, int nSE, int nIL, const char *const *files, int nfiles)
{
    (void)nIL;
    if (num_chunks <= 0 || nfiles <= 0) return fail(-1, "bad arguments");
    int numPEFiles = nPE * 2, numSEFiles = nSE;
    uint64_t *fsize = (uint64_t *)calloc((size_t)nfiles, sizeof(uint64_t));
    uint64_t total = 0;
    for (int i = 0; i < nfiles; i++) {
        if (file_size(files[i], &fsize[i])) { free(fsize); return fail(-1, "cannot stat %s", files[i]); }
        total += fsize[i];
    }
    uint64_t per = total / (uint64_t)num_chunks;
    char path[4096];
    snprintf(path, sizeof(path), "%s.idx", files[0]);
    FILE *out = fopen(path, "wb");
    if (!out) { free(fsize); return fail(-1, "cannot create %s", path); }
    orc_fileindex *fI = (orc_fileindex *)calloc(1, sizeof(orc_fileindex));
    char *chunk = (char *)calloc(5000 + 64, 1);
    int rc = 0, interleaved = 0;
    int maxReadSize = 300;
    fI->fileNo = 0;
    fI->offset = 0;
    for (int i = 0; i < num_chunks; i++) {
        int flag = 0;
        if ((rc = get_file_index(i, per, nfiles, fsize, numPEFiles, &fI->fileNo, &fI->offset))) break;
        if (fI->fileNo >= numPEFiles + numSEFiles) interleaved = 1;
        if (fsize[fI->fileNo] - fI->offset < (uint64_t)(int64_t)maxReadSize) { /* 219-225 */
            if (fI->fileNo + 1 == nfiles) break;
            fI->fileNo++;
            fI->offset = 0;
        }
        if (i == 0) {
            fI->offset2 = 0;
            fI->readId = 0;
            fwrite(fI, sizeof(orc_fileindex), 1, out);
            continue;
        }
        FILE *fp = fopen(files[fI->fileNo], "rb");
        if (!fp) { rc = fail(-1, "cannot open %s", files[fI->fileNo]); break; }
        fseek(fp, (long)fI->offset, SEEK_SET);
        memset(chunk, 0, 5000);
        size_t nread = fread(chunk, 1, 3000, fp);
        fclose(fp);
        if (!(nread > (size_t)(int64_t)maxReadSize)) { rc = fail(-1, "chunk %d: fread %zu <= maxReadSize %d", i, nread, maxReadSize); break; }
        int k = 0;
        size_t j;
        const size_t LIM = 4990;
#define BOUND(x) do { if ((x) >= LIM) { rc = fail(-1, "chunk %d: record longer than the 3000-byte probe", i); goto done; } } while (0)
        for (j = 0; j < nread; j++) {
            if (chunk[j] == '\n' && chunk[j + 1] == '+') { /* 260-267 */
                flag = 1;
                j = j + 2;
                while (chunk[j] != '\n') { j++; BOUND(j); }
                j++;
            }
            k = 0;
            if (flag && chunk[j] == '\n') { /* 269-301 */
                j++;
                k++;
                if (interleaved) {
                    char rName[256], rName2[256];
                    size_t kk;
                    for (kk = j; chunk[kk] != '/' && chunk[kk] != ' '; kk++) { BOUND(kk); if (kk - j < 255) rName[kk - j] = chunk[kk]; }
                    rName[(kk - j) < 255 ? kk - j : 255] = '\0';
                    for (int l = 0; l < 4; l++) { while (chunk[kk] != '\n') { kk++; BOUND(kk); } kk++; }
                    size_t j2 = kk;
                    while (chunk[kk] != ' ' && chunk[kk] != '/') { BOUND(kk); if (kk - j2 < 255) rName2[kk - j2] = chunk[kk]; kk++; }
                    rName2[(kk - j2) < 255 ? kk - j2 : 255] = '\0';
                    if (strcmp(rName, rName2)) j = j2;
                    k = (int)kk;
                    break;
                } else
                    break;
            }
        }
        if (2 * k + 30 < maxReadSize) maxReadSize = 2 * k + 30; /* 303 */
        if (flag != 1) { rc = fail(-1, "chunk %d: no record boundary found", i); break; }
        fI->offset = fI->offset + j;
        if (fI->fileNo < numPEFiles) {
            char name[256];
            int n = 0;
            while (chunk[j] != '\n' && n < 255) name[n++] = chunk[j++];
            if (n < 2) { rc = fail(-1, "short read name"); break; }
            name[n - 2] = '\0';
            if ((rc = find_ofs_other_end(files, fsize, fI->fileNo, fI->offset, name, &fI->offset2))) break;
        } else
            fI->offset2 = 0;
        fI->readId = 0;
        fwrite(fI, sizeof(orc_fileindex), 1, out);
    }
#undef BOUND
done:
    if (!rc) { /* sentinel 338-361 */
        if (nfiles == numPEFiles) {
            fI->fileNo = nfiles - 2;
            fI->offset = fsize[nfiles - 2];
            fI->offset2 = fsize[nfiles - 1];
        } else {
            fI->fileNo = nfiles - 1;
            fI->offset = fsize[nfiles - 1];
            fI->offset2 = 0;
        }
        fI->readId = 0;
        fwrite(fI, sizeof(orc_fileindex), 1, out);
    }
    fclose(out);
    free(chunk);
    free(fI);
    free(fsize);
    return rc;
}

/* ======================================================================
 * k-mer arithmetic (countMmers.c / enumerateKmers.c), k <= 32
 * ====================================================================== */
static const int CMAP[4] = {0, 1, 3, 2};  /* countMmers.c:85 */
static const int RCMAP[4] = {3, 2, 0, 1}; /* countMmers.c:86 */
static inline uint64_t kmask(int K) { return 2 * K >= 64 ? 0ull : (1ull << (2 * K)) - 1; } /* x86 shift semantics at K=32 */
/* ORC_MUTANT (0 here) builds a copy with one planted mistake, which the golden
 * pins must catch (tests/test_oracle_mutants.py): 1 the pair step's size
 * gate asking both components to be small, 2 the top bit of the next character's class dropped from
 * the adj byte, 3 the filter's ratio at 0.25, 4 a 9-base m-mer prefix */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
static inline uint64_t mmer_prefix(uint64_t kL, int K) { return kL >> ((K - (ORC_MUTANT == 4 ? 9 : 10)) << 1); } /* countMmers.c:60-81, m = 10 */
static inline char rcchar(char c) { return "TGAC   N"[(c >> 1) & 7]; } /* enumerateKmers.c:237 */
static inline uint8_t adj_fwd(char prev, char next)
{
    if (ORC_MUTANT == 2) return (uint8_t)(((prev << 2) & 0xF8) | ((next >> 1) & 0x3));
    return (uint8_t)(((prev << 2) & 0xF8) | ((next >> 1) & 0x7));
}

/* plain counter: countMmers.c:83-217 */
static int count_plain(const char *s, int L, int K, uint32_t *counts)
{
    if (L <= K) return 0;
    int last = L - K, lastN = -1, n = 0;
    uint64_t F = 0, R = 0, mask = kmask(K);
    for (int j = 0; j < K; j++) {
        if (s[j] == 'N') lastN = j;
        R |= (uint64_t)RCMAP[(s[j] & 6) >> 1] << (j << 1);
        F = (F << 2) | (uint64_t)CMAP[(s[j] & 6) >> 1];
    }
    if (lastN == -1) {
        uint64_t b = mmer_prefix(F < R ? F : R, K);
        if (counts) counts[b]++;
        n++;
    }
    for (int i = 1; i <= last; i++) {
        char c = s[i + K - 1];
        R = (R >> 2) | ((uint64_t)RCMAP[(c & 6) >> 1] << ((K - 1) << 1));
        F = ((F << 2) | (uint64_t)CMAP[(c & 6) >> 1]) & mask;
        if (c == 'N') lastN = i + K - 1;
        if (i > lastN) {
            uint64_t b = mmer_prefix(F < R ? F : R, K);
            if (counts) counts[b]++;
            n++;
        }
    }
    return n;
}

/* 3-minimizer counter: countMmers_3minimizers.c:83-271 */
static int count_min3(const char *s, int L, int K, uint32_t *counts)
{
    if (L <= K) return 0;
    int last = L - K, lastN = -1, n = 0;
    uint64_t F = 0, R = 0, mask = kmask(K);
    uint64_t minK = 0, minB = 0;
    int next = last / 3;
    for (int j = 0; j < K; j++) {
        if (s[j] == 'N') lastN = j;
        R |= (uint64_t)RCMAP[(s[j] & 6) >> 1] << (j << 1);
        F = (F << 2) | (uint64_t)CMAP[(s[j] & 6) >> 1];
    }
    if (lastN == -1) {
        uint64_t c = F < R ? F : R;
        minK = c;
        minB = mmer_prefix(c, K);
    }
    for (int i = 1; i <= last; i++) {
        char ch = s[i + K - 1];
        R = (R >> 2) | ((uint64_t)RCMAP[(ch & 6) >> 1] << ((K - 1) << 1));
        F = ((F << 2) | (uint64_t)CMAP[(ch & 6) >> 1]) & mask;
        if (ch == 'N') lastN = i + K - 1;
        if (i > lastN) {
            uint64_t c = F < R ? F : R;
            uint64_t b = mmer_prefix(c, K);
            if (i > next) {
                if (minK != 0) { if (counts) counts[minB]++; n++; }
                minK = c;
                minB = b;
                next += last / 3;
                if (next + last / 3 > last) next = last;
            } else if (c < minK || minK == 0) {
                minK = c;
                minB = b;
            }
        }
    }
    if (minK != 0) { if (counts) counts[minB]++; n++; }
    return n;
}

int orc_count_read(const char *seq, int len, int K, int min3, uint32_t *counts)
{
    return min3 ? count_min3(seq, len, K, counts) : count_plain(seq, len, K, counts);
}

static inline void put_tuple(uint8_t *t, uint64_t kmer, uint8_t adj, uint64_t label)
{
    memcpy(t, &kmer, 8);
    t[8] = adj;
    t[9] = (uint8_t)(label / U32MAX); /* enumerateKmers.c:360: (uint16)(id/UINT32_MAX) << 8 */
    uint32_t rem = (uint32_t)(label % U32MAX);
    memcpy(t + 10, &rem, 4);
}
static inline uint64_t t_kmer(const uint8_t *t) { uint64_t k; memcpy(&k, t, 8); return k; }
static inline uint64_t t_label(const uint8_t *t)
{
    uint32_t rem;
    memcpy(&rem, t + 10, 4);
    return (uint64_t)rem + (uint64_t)t[9] * U32MAX;
}

/* plain enumerator: enumerateKmers.c:127-371 (P = 1, so lBuf order is output order) */
static int enum_plain(const char *s, int L, int K, uint64_t lo, uint64_t hi, uint64_t label, uint8_t *out, int cap)
{
    if (L <= K) return 0;
    int last = L - K, lastN = -1, n = 0;
    if (last >= 300) return -1; /* enumerateKmers.c:155 */
    uint64_t F = 0, R = 0, mask = kmask(K);
    char prev = 'N', next = s[K];
    for (int j = 0; j < K; j++) {
        if (s[j] == 'N') lastN = j;
        R |= (uint64_t)RCMAP[(s[j] & 6) >> 1] << (j << 1);
        F = (F << 2) | (uint64_t)CMAP[(s[j] & 6) >> 1];
    }
#define EMIT()                                                                        \
    do {                                                                              \
        if (F < R) {                                                                  \
            uint64_t b = mmer_prefix(F, K);                                           \
            if (b >= lo && b < hi) {                                                  \
                if (n >= cap) return -2;                                              \
                put_tuple(out + 14 * (size_t)n++, F, adj_fwd(prev, next), label);     \
            }                                                                         \
        } else {                                                                      \
            uint64_t b = mmer_prefix(R, K);                                           \
            if (b >= lo && b < hi) {                                                  \
                if (n >= cap) return -2;                                              \
                char p2 = rcchar(prev), n2 = rcchar(next);                            \
                put_tuple(out + 14 * (size_t)n++, R, adj_fwd(n2, p2), label);         \
            }                                                                         \
        }                                                                             \
    } while (0)
    if (lastN == -1) EMIT();
    for (int i = 1; i <= last; i++) {
        prev = s[i - 1];
        next = i == last ? 'N' : s[i + K];
        char c = s[i + K - 1];
        R = (R >> 2) | ((uint64_t)RCMAP[(c & 6) >> 1] << ((K - 1) << 1));
        F = ((F << 2) | (uint64_t)CMAP[(c & 6) >> 1]) & mask;
        if (c == 'N') lastN = i + K - 1;
        if (i > lastN) EMIT();
    }
#undef EMIT
    return n;
}

/* 3-minimizer enumerator: enumerateKmers_3minimizers.c:129-474 */
static int enum_min3(const char *s, int L, int K, uint64_t lo, uint64_t hi, uint64_t label, uint8_t *out, int cap)
{
    if (L <= K) return 0;
    int last = L - K, lastN = -1, n = 0;
    if (last >= 300) return -1;
    uint64_t F = 0, R = 0, mask = kmask(K);
    uint64_t minK = 0;
    char minPrev = 'N', minNext = 'N';
    char prev = 'N', next = s[K];
    for (int j = 0; j < K; j++) {
        if (s[j] == 'N') lastN = j;
        R |= (uint64_t)RCMAP[(s[j] & 6) >> 1] << (j << 1);
        F = (F << 2) | (uint64_t)CMAP[(s[j] & 6) >> 1];
    }
    if (lastN == -1) {
        if (F < R) {
            uint64_t b = mmer_prefix(F, K);
            if (b >= lo && b < hi) { minK = F; minPrev = prev; minNext = next; }
        } else {
            uint64_t b = mmer_prefix(R, K);
            if (b >= lo && b < hi) { minK = R; minPrev = rcchar(prev); minNext = rcchar(next); }
        }
    }
    int nextW = last / 3;
    for (int i = 1; i <= last; i++) {
        prev = s[i - 1];
        next = i == last ? 'N' : s[i + K];
        char c = s[i + K - 1];
        R = (R >> 2) | ((uint64_t)RCMAP[(c & 6) >> 1] << ((K - 1) << 1));
        F = ((F << 2) | (uint64_t)CMAP[(c & 6) >> 1]) & mask;
        if (c == 'N') lastN = i + K - 1;
        if (i <= lastN) continue;
        int fwd = F < R;
        uint64_t cur = fwd ? F : R;
        uint64_t b = mmer_prefix(cur, K);
        if (!(b >= lo && b < hi)) continue;
        char cp = fwd ? prev : rcchar(prev), cn = fwd ? next : rcchar(next);
        if (i > nextW) {
            if (minK != 0) {
                if (n >= cap) return -2;
                /* 328 (forward trigger) / 388 (reverse trigger) */
                uint8_t a = fwd ? adj_fwd(minPrev, minNext) : adj_fwd(minNext, minPrev);
                put_tuple(out + 14 * (size_t)n++, minK, a, label);
            }
            minK = cur; minPrev = cp; minNext = cn;
            nextW += last / 3;
            if (nextW + last / 3 > last) nextW = last;
        } else if (cur < minK || minK == 0) {
            minK = cur; minPrev = cp; minNext = cn;
        }
    }
    if (minK != 0) { /* 418-426 */
        if (n >= cap) return -2;
        put_tuple(out + 14 * (size_t)n++, minK, adj_fwd(minNext, minPrev), label);
    }
    return n;
}

int orc_enum_read(const char *seq, int len, int K, int min3, uint64_t lo, uint64_t hi, uint64_t label,
                  uint8_t *out, int cap)
{
    return min3 ? enum_min3(seq, len, K, lo, hi, label, out, cap) : enum_plain(seq, len, K, lo, hi, label, out, cap);
}

/* ======================================================================
 * genHistograms (genHistograms.c)
 * ====================================================================== */

/* read the bytes of chunk [rec[i], rec[i+1]) across files: genHistograms.c:207-251 */
static int load_chunk_bytes(const char *const *files, const uint64_t *fsize, const orc_fileindex *a,
                            const orc_fileindex *b, char **buf, uint64_t *len)
{
    uint64_t total = 0;
    for (int j = a->fileNo; j <= b->fileNo; j++) {
        uint64_t so = j == a->fileNo ? a->offset : 0;
        uint64_t eo = j == b->fileNo ? b->offset : fsize[j];
        total += eo - so;
    }
    char *p = (char *)malloc(total + 16);
    if (!p) return fail(-2, "oom");
    int f = a->fileNo;
    FILE *fp = fopen(files[f], "rb");
    if (!fp) { free(p); return fail(-1, "cannot open %s", files[f]); }
    fseek(fp, (long)a->offset, SEEK_SET);
    uint64_t got = 0;
    while (got < total) {
        got += fread(p + got, 1, total - got, fp);
        if (got < total) {
            fclose(fp);
            f++;
            fp = fopen(files[f], "rb");
            if (!fp) { free(p); return fail(-1, "cannot open %s", files[f]); }
        }
    }
    fclose(fp);
    memset(p + total, 0, 16);
    *buf = p;
    *len = total;
    return 0;
}

int orc_gen_histograms(int K, int nPE, int nSE, int nIL, const char *const *files, int nfiles, int min3)
{
    (void)nIL; (void)nSE;
    if (K <= 10 || K > 32) return fail(-1, "K must be in [11, 32]");
    if (nPE > 0) return fail(-1, "two-file paired-end input is not supported");
    uint64_t *fsize = (uint64_t *)calloc((size_t)nfiles, sizeof(uint64_t));
    for (int i = 0; i < nfiles; i++)
        if (file_size(files[i], &fsize[i])) { free(fsize); return fail(-1, "cannot stat %s", files[i]); }
    char path[4096];
    snprintf(path, sizeof(path), "%s.idx", files[0]);
    uint64_t isz;
    if (file_size(path, &isz) || isz < 2 * REC_SIZE || isz % REC_SIZE) { free(fsize); return fail(-1, "bad index %s", path); }
    int numRecs = (int)(isz / REC_SIZE) - 1;
    orc_fileindex *fI = (orc_fileindex *)malloc((size_t)(numRecs + 1) * REC_SIZE);
    uint64_t *hist = (uint64_t *)calloc(ORC_NBUCKETS + 2, sizeof(uint64_t));
    uint64_t *readCounts = (uint64_t *)calloc((size_t)numRecs + 2, sizeof(uint64_t));
    if (!fI || !hist || !readCounts) { free(fsize); return fail(-2, "oom"); }
    FILE *fp = fopen(path, "rb");
    size_t got = fread(fI, REC_SIZE, (size_t)numRecs + 1, fp);
    fclose(fp);
    if (got != (size_t)numRecs + 1) { free(fsize); free(fI); free(hist); free(readCounts); return fail(-1, "short index"); }
    uint64_t totalKmers = 0;
    int rc = 0;
    for (int i = 0; i < numRecs && !rc; i++) { /* 200-265 */
        memset(fI[i].kmerFreqCount, 0, sizeof(fI[i].kmerFreqCount));
        char *buf;
        uint64_t len;
        if ((rc = load_chunk_bytes(files, fsize, &fI[i], &fI[i + 1], &buf, &len))) break;
        uint64_t ind = 0;
        while (ind < len) { /* getNextRead 60-80 */
            uint64_t cur = ind;
            while (cur < len && buf[cur] != '\n') cur++;
            cur++;
            uint64_t seq = cur;
            int L = 0;
            while (cur < len && buf[cur] != '\n') { cur++; L++; }
            if (cur >= len) { rc = fail(-1, "truncated record in chunk %d", i); break; }
            cur++; /* skip the newline that ends the sequence, then the '+' line */
            while (cur < len && buf[cur] != '\n') cur++;
            cur++;
            cur += (uint64_t)L + 1;
            ind = cur;
            int n = min3 ? count_min3(buf + seq, L, K, fI[i].kmerFreqCount)
                         : count_plain(buf + seq, L, K, fI[i].kmerFreqCount);
            totalKmers += (uint64_t)n;
            readCounts[i + 2]++;
        }
        free(buf);
    }
    if (!rc) {
        for (int i = 0; i < numRecs; i++)
            for (uint32_t b = 0; b < ORC_NBUCKETS; b++) hist[b + 2] += fI[i].kmerFreqCount[b];
        uint64_t readcount = 0; /* 280-291 */
        for (int i = 1; i <= numRecs; i++) {
            readcount += readCounts[i + 1];
            fI[i].readId = (uint32_t)(readcount % U32MAX);
            if (readcount / U32MAX >= 256) { rc = fail(-1, "too many reads"); break; }
            fI[i].readIdExtra = (uint8_t)(readcount / U32MAX);
        }
    }
    if (!rc) {
        for (uint32_t i = 1; i <= ORC_NBUCKETS; i++) hist[i] = hist[i - 1] + hist[i + 1]; /* 325-326 */
        if (hist[ORC_NBUCKETS] != totalKmers) rc = fail(-1, "histogram total mismatch");
    }
    if (!rc) {
        snprintf(path, sizeof(path), min3 ? "%s.KmerIdx_3min%d" : "%s.KmerIdx%d", files[0], K);
        FILE *o = fopen(path, "wb");
        if (!o) rc = fail(-1, "cannot create %s", path);
        else { fwrite(fI, REC_SIZE, (size_t)numRecs + 1, o); fclose(o); }
        snprintf(path, sizeof(path), min3 ? "%s.kmerHist_3min%d" : "%s.kmerHist%d", files[0], K);
        o = fopen(path, "wb");
        if (!o) rc = fail(-1, "cannot create %s", path);
        else { fwrite(hist, sizeof(uint64_t), ORC_NBUCKETS + 1, o); fclose(o); }
    }
    free(fsize); free(fI); free(hist); free(readCounts);
    return rc;
}

/* ======================================================================
 * metapart (metapart.c)
 * ====================================================================== */

/* metapart.c:3720-3791 with numTasks = 1 (process ranges equal scan ranges) */
int orc_plan_ranges(const uint64_t *H, int S, int T, uint64_t *scan, uint64_t *thread)
{
    if (S <= 0 || T <= 0) return fail(-1, "bad S/T");
    uint64_t tot = H[ORC_NBUCKETS];
    uint64_t avgS = tot / (uint64_t)S;
    if (tot % (uint64_t)S) avgS++;
    memset(scan, 0, sizeof(uint64_t) * (size_t)(S + 1));
    int r = 1;
    uint32_t i;
    for (i = 0; i < ORC_NBUCKETS; i++)
        if (H[i] >= avgS * (uint64_t)r) {
            if (r > S) return fail(-3, "scan range walk overshoots the array (input too small for S)");
            scan[r] = i;
            r++;
        }
    if (r == S) scan[r] = i;
    /* process ranges: P = 1 gives proc[r] = scan[r] at identical thresholds */
    uint64_t avgT = avgS / (uint64_t)T;
    memset(thread, 0, sizeof(uint64_t) * (size_t)(S * T + 1));
    r = 1;
    for (i = 0; i < ORC_NBUCKETS; i++)
        if (H[i] >= (uint64_t)r * avgT) {
            if (r > S * T) return fail(-3, "thread range walk overshoots the array (input too small for T)");
            thread[r] = (r % T == 0) ? scan[r / T] : i;
            r++;
        }
    if (r == S * T) thread[r] = i;
    return 0;
}

typedef struct {
    /* inputs */
    const char *const *files;
    int nfiles, K, S, T, min3;
    uint64_t bufsz;
    uint64_t *fsize;
    int numPEFiles, numSEFiles;
    int C;
    orc_fileindex *fI; /* C+1 records */
    uint64_t N;
    uint64_t *H;
    uint64_t *scan, *thread;
    uint64_t *tkc;  /* threadKmerCounts T x 2^20 */
    uint64_t *sendThreadSizes; /* S x T */
    /* state */
    uint64_t *parent; /* ORC_BAD or parent id */
    uint8_t *filtered;
    uint32_t *sizes;
    /* dump */
    int dump_scan;
    uint8_t *dump_tuples;
    uint64_t dump_n;
    uint64_t *dump_parent;
} mp_t;

static int g_dump_scan = -1;
static uint8_t *g_dump_tuples = NULL;
static uint64_t g_dump_n = 0, g_dump_nreads = 0;
static uint64_t *g_dump_parent = NULL;
void orc_set_dump_scan(int scan) { g_dump_scan = scan; }
/* per scan: runs the filter evaluated (each k-mer change inside a segment
 * ends one, metapart.c:2288-2290) and runs it zeroed (both sides branch) */
#define ORC_MAX_SCANS 256
static uint64_t g_runs_eval[ORC_MAX_SCANS], g_runs_zero[ORC_MAX_SCANS];
int orc_get_filter_counts(int scan, uint64_t *evaluated, uint64_t *zeroed)
{
    if (scan < 0 || scan >= ORC_MAX_SCANS) return -1;
    *evaluated = g_runs_eval[scan];
    *zeroed = g_runs_zero[scan];
    return 0;
}
int orc_get_dump(uint8_t **tuples, uint64_t *n, uint64_t **parent_after, uint64_t *nreads)
{
    *tuples = g_dump_tuples; *n = g_dump_n; *parent_after = g_dump_parent; *nreads = g_dump_nreads;
    g_dump_tuples = NULL; g_dump_parent = NULL;
    return 0;
}

/* loadChunk (metapart.c:371-500) for SE / interleaved input: bytes of chunk c
 * and, sticky per thread, the interleave flag and start offset. */
static int mp_load_chunk(mp_t *m, int c, char **buf, uint64_t *len, int *interleave, uint64_t *ilStart)
{
    int cur = m->fI[c].fileNo, endF = m->fI[c + 1].fileNo;
    uint64_t curOfs = m->fI[c].offset, endOfs = m->fI[c + 1].offset;
    uint64_t total = 0;
    {
        int f = cur;
        uint64_t o = curOfs;
        while (f < endF || (f == endF && o < endOfs)) {
            total += (f < endF) ? m->fsize[f] - o : endOfs - o;
            f++;
            o = 0;
        }
    }
    char *p = (char *)malloc(total + 16);
    if (!p) return fail(-2, "oom");
    uint64_t bo = 0, remaining = m->bufsz;
    while (cur < endF || (cur == endF && curOfs < endOfs)) {
        if (cur >= m->numPEFiles + m->numSEFiles && *interleave != 1) { *interleave = 1; *ilStart = bo; }
        uint64_t want = (cur < endF) ? m->fsize[cur] - curOfs : endOfs - curOfs;
        if (!(remaining > want)) { free(p); return fail(-1, "chunk %d larger than fastqBufferSize", c); } /* 471 */
        FILE *fp = fopen(m->files[cur], "rb");
        if (!fp) { free(p); return fail(-1, "cannot open %s", m->files[cur]); }
        fseek(fp, (long)curOfs, SEEK_SET);
        uint64_t got = fread(p + bo, 1, want, fp);
        fclose(fp);
        if (got != want) { free(p); return fail(-1, "short read"); }
        remaining -= want;
        bo += want;
        cur++;
        curOfs = 0;
    }
    memset(p + total, 0, 16);
    *buf = p;
    *len = total;
    return 0;
}

/* getNextRead (metapart.c:296-369), SE / interleaved branch.  Returns 1 or 2
 * (number of reads), -1 at end; seq pointers and lengths out. */
static int next_read(const char *b, uint64_t size, uint64_t *ind, int interleave, uint64_t ilStart,
                     uint64_t *s1, int *l1, uint64_t *s2, int *l2)
{
    *l1 = *l2 = 0;
    if (*ind >= size) return -1;
    *s1 = *ind;
    while (*ind < size && b[*ind] != '\n') { (*ind)++; (*l1)++; }
    (*ind)++;
    while (*ind < size && b[*ind] != '\n') (*ind)++;
    (*ind)++;
    *ind += (uint64_t)*l1;
    (*ind)++;
    int ret = 1;
    if (*ind < size) {
        while (*ind < size && b[*ind] != '\n') (*ind)++;
        (*ind)++;
        if (interleave && *ind >= ilStart) {
            *s2 = *ind;
            while (*ind < size && b[*ind] != '\n') { (*ind)++; (*l2)++; }
            (*ind)++;
            while (*ind < size && b[*ind] != '\n') (*ind)++;
            (*ind)++;
            *ind += (uint64_t)*l2;
            (*ind)++;
            ret = 2;
            if (*ind < size) {
                while (*ind < size && b[*ind] != '\n') (*ind)++;
                (*ind)++;
            }
        }
    }
    return ret;
}

/* getNextFullRead (metapart.c:211-294), SE / interleaved branch */
static int next_full_read(const char *b, uint64_t size, uint64_t *ind, int interleave, uint64_t ilStart,
                          uint64_t *r1, uint64_t *len1, uint64_t *r2, uint64_t *len2)
{
    *len1 = *len2 = 0;
    if (*ind >= size) return -1;
    *r1 = *ind;
    uint64_t old = *ind;
    uint64_t L = 0;
    while (*ind < size && b[*ind] != '\n') (*ind)++;
    (*ind)++;
    while (*ind < size && b[*ind] != '\n') { (*ind)++; L++; }
    (*ind)++;
    while (*ind < size && b[*ind] != '\n') (*ind)++;
    (*ind)++;
    *ind += L;
    (*ind)++;
    *len1 = *ind - old;
    if (interleave && *ind >= ilStart) {
        old = *ind;
        *r2 = *ind;
        L = 0;
        while (*ind < size && b[*ind] != '\n') (*ind)++;
        (*ind)++;
        while (*ind < size && b[*ind] != '\n') { (*ind)++; L++; }
        (*ind)++;
        while (*ind < size && b[*ind] != '\n') (*ind)++;
        (*ind)++;
        *ind += L;
        (*ind)++;
        *len2 = *ind - old;
        return 2;
    }
    return 1;
}

/* unionfind.h:19-55 on a u64 parent array with the BAD sentinel */
static uint64_t find_root(uint64_t *parent, uint64_t e)
{
    uint64_t p = parent[e];
    if (p == ORC_BAD) return e;
    while (p != e) {
        if (p == ORC_BAD) break;
        uint64_t pp = parent[p];
        if (pp != ORC_BAD) { parent[e] = pp; p = pp; }
        if (p != ORC_BAD) e = p; else break;
        p = parent[e];
    }
    return e;
}

/* one radixPass (metapart.c:1494-1598): stable counting pass on tuple byte `byte` */
static void radix_pass(const uint8_t *in, uint8_t *out, uint64_t start, uint64_t n, int byte)
{
    uint64_t cnt[257];
    memset(cnt, 0, sizeof(cnt));
    for (uint64_t i = start; i < start + n; i++) cnt[in[14 * i + byte] + 1]++;
    for (int d = 1; d <= 256; d++) cnt[d] += cnt[d - 1];
    for (uint64_t i = start; i < start + n; i++) {
        uint64_t pos = start + cnt[in[14 * i + byte]]++;
        memcpy(out + 14 * pos, in + 14 * i, 14);
    }
}

/* generateKmersInScan (metapart.c:1192-1476) with getReadAndKmerCount (501-573):
 * tuples of scan s in generation order, thread by thread. */
static int mp_enumerate(mp_t *m, int s, uint8_t **tuples_out, uint64_t *n_out, uint64_t *per_thread)
{
    uint64_t total = 0;
    for (int t = 0; t < m->T; t++) total += m->sendThreadSizes[(size_t)s * m->T + t];
    uint8_t *tp = (uint8_t *)malloc(14 * (total + 1));
    if (!tp) return fail(-2, "oom");
    uint64_t lo = m->scan[s], hi = m->scan[s + 1];
    uint64_t n = 0;
    uint8_t tmp[14 * 300];
    int rc = 0;
    int cpt = m->C / m->T;
    for (int t = 0; t < m->T && !rc; t++) {
        uint64_t t0 = n;
        int interleave = 0;
        uint64_t ilStart = 0;
        for (int c = t * cpt; c < (t + 1) * cpt && !rc; c++) {
            uint64_t id = (uint64_t)m->fI[c].readIdExtra * U32MAX + m->fI[c].readId;
            char *b;
            uint64_t len;
            if ((rc = mp_load_chunk(m, c, &b, &len, &interleave, &ilStart))) break;
            uint64_t ind = 0;
            while (ind < len && b[ind] != '\n') ind++; /* 1326-1327: skip the first header */
            ind++;
            if (!(ind < len)) { free(b); rc = fail(-1, "empty chunk %d", c); break; }
            uint64_t s1, s2;
            int l1, l2, ret;
            while ((ret = next_read(b, len, &ind, interleave, ilStart, &s1, &l1, &s2, &l2)) != -1) {
                for (int mate = 0; mate < ret; mate++) {
                    uint64_t rid = id + (uint64_t)mate;
                    uint64_t label = m->parent == NULL ? rid : find_root(m->parent, rid); /* 518-557 */
                    const char *sq = b + (mate ? s2 : s1);
                    int L = mate ? l2 : l1;
                    int k = m->min3 ? enum_min3(sq, L, m->K, lo, hi, label, tmp, 300)
                                    : enum_plain(sq, L, m->K, lo, hi, label, tmp, 300);
                    if (k < 0) { rc = fail(-1, "read longer than 299+K bases"); break; }
                    if (n + (uint64_t)k > total) { rc = fail(-4, "scan %d: more tuples than the histogram sized (3-minimizer windows at S>1?)", s); break; }
                    memcpy(tp + 14 * n, tmp, 14 * (size_t)k);
                    n += (uint64_t)k;
                }
                if (rc) break;
                id += (uint64_t)ret;
            }
            free(b);
        }
        if (!rc && n - t0 != m->sendThreadSizes[(size_t)s * m->T + t])
            rc = fail(-4, "scan %d thread %d: %llu tuples, histogram says %llu", s, t, (unsigned long long)(n - t0),
                      (unsigned long long)m->sendThreadSizes[(size_t)s * m->T + t]);
        per_thread[t] = n - t0;
    }
    if (rc) { free(tp); return rc; }
    *tuples_out = tp;
    *n_out = n;
    return 0;
}

/* sortKmerList (metapart.c:1852-2267): range partition by thread range, then
 * 9 stable byte passes (bytes 8, 0..7) per segment. */
static int mp_sort(mp_t *m, int s, uint8_t *in, uint64_t n, const uint64_t *per_thread, uint8_t **sorted_out,
                   uint64_t *lrc, uint64_t *roff)
{
    int T = m->T;
    const uint64_t *rng = &m->thread[(size_t)s * T];
    /* 1930-1949: localRangeCounts / rangeOffsets, one thread step per bucket */
    memset(lrc, 0, sizeof(uint64_t) * (size_t)(T + 1));
    memset(roff, 0, sizeof(uint64_t) * (size_t)T);
    uint64_t *lro = (uint64_t *)calloc((size_t)T, sizeof(uint64_t));
    for (int f = 0; f < T; f++) {
        int ct = 0;
        for (uint64_t b = rng[0]; b < rng[T]; b++) {
            if (b >= rng[ct + 1]) {
                ct++;
                if (ct >= T) { free(lro); return fail(-3, "degenerate thread ranges"); }
                lro[ct] = lro[ct - 1] + lrc[ct - 1];
                roff[ct] = lro[ct];
            }
            lrc[ct] += m->tkc[(size_t)f * ORC_NBUCKETS + b];
        }
    }
    free(lro);
    /* calcRangeCounts 1810-1850 and the column-major offsets 1961-1969 */
    uint64_t *trc = (uint64_t *)calloc((size_t)T * T, sizeof(uint64_t));
    uint64_t *tro = (uint64_t *)calloc((size_t)T * T, sizeof(uint64_t));
    for (int tid = 0; tid < T; tid++)
        for (int t = 0; t < T; t++)
            for (uint64_t p = rng[t]; p < rng[t + 1]; p++) trc[(size_t)tid * T + t] += m->tkc[(size_t)tid * ORC_NBUCKETS + p];
    uint64_t cur = 0;
    for (int i = 0; i < T; i++)
        for (int f = 0; f < T; f++) { tro[(size_t)f * T + i] = cur; cur += trc[(size_t)f * T + i]; }
    /* the two offset computations must describe the same layout */
    for (int t = 0; t < T; t++) {
        uint64_t binTotal = 0;
        for (int f = 0; f < T; f++) binTotal += trc[(size_t)f * T + t];
        if (binTotal != lrc[t] || (t > 0 && roff[t] != tro[t])) {
            free(trc); free(tro);
            return fail(-3, "degenerate thread ranges (empty or unordered segment)");
        }
    }
    uint8_t *sorted = (uint8_t *)malloc(14 * (n + 1));
    uint8_t *tmp = (uint8_t *)malloc(14 * (n + 1));
    if (!sorted || !tmp) { free(trc); free(tro); free(sorted); free(tmp); return fail(-2, "oom"); }
    /* 1971-2111: range partition */
    uint64_t g = 0;
    for (int tid = 0; tid < T; tid++) {
        uint64_t *ptr = &tro[(size_t)tid * T];
        for (uint64_t j = 0; j < per_thread[tid]; j++, g++) {
            uint64_t b = mmer_prefix(t_kmer(in + 14 * g), m->K);
            int bin = 0;
            while (bin <= T && b >= rng[bin]) bin++;
            bin--;
            if (bin < 0 || bin >= T) { free(trc); free(tro); free(sorted); free(tmp); return fail(-3, "tuple outside the thread ranges"); }
            memcpy(sorted + 14 * ptr[bin]++, in + 14 * g, 14);
        }
    }
    free(trc); free(tro);
    /* 2167-2198: passes 5..13 = bytes 8,0,1,...,7; ping-pong then copy back */
    static const int bytes[9] = {8, 0, 1, 2, 3, 4, 5, 6, 7};
    for (int t = 0; t < T; t++) {
        uint8_t *a = sorted, *bb = tmp;
        for (int p = 0; p < 9; p++) {
            radix_pass(a, bb, roff[t], lrc[t], bytes[p]);
            uint8_t *x = a; a = bb; bb = x;
        }
        /* after 9 passes the data is in tmp; copy segment back */
        memcpy(sorted + 14 * roff[t], tmp + 14 * roff[t], 14 * lrc[t]);
    }
    free(tmp);
    *sorted_out = sorted;
    return 0;
}

/* markMischiefKmers (metapart.c:2270-2519) */
static void mp_filter(mp_t *m, int scan, uint8_t *st, const uint64_t *lrc, const uint64_t *roff)
{
    uint64_t nev = 0, nz = 0;
    for (int seg = 0; seg < m->T; seg++) {
        uint64_t first = roff[seg];
        uint64_t curK = t_kmer(st + 14 * first), curOfs = first;
        int before[8] = {0}, after[8] = {0}, deg = 0;
        for (uint64_t j = roff[seg]; j < roff[seg] + lrc[seg]; j++) {
            uint64_t nk = t_kmer(st + 14 * j);
            if (curK != nk) {
                int nIn = 0, nOut = 0, maxIn = 0, maxOut = 0, mInC = 0, mOutC = 0, sInC = 0, sOutC = 0;
                for (int k = 0; k < 7; k++) {
                    if (before[k]) {
                        if (before[k] > mInC) { sInC = mInC; mInC = before[k]; maxIn = k; }
                        else if (before[k] > sInC) sInC = before[k];
                        nIn++;
                    }
                    if (after[k]) {
                        if (after[k] > mOutC) { sOutC = mOutC; mOutC = after[k]; maxOut = k; }
                        else if (after[k] > sOutC) sOutC = after[k];
                        nOut++;
                    }
                }
                before[7] = after[7] = 0;
                nev++;
                nz += nIn > 1 && nOut > 1;
                if (nIn > 1 || nOut > 1) {
                    float sInR = (float)sInC / deg; /* fp32 divide, compared as double (2423) */
                    float sOutR = (float)sOutC / deg;
                    for (uint64_t k = curOfs; k < j; k++) {
                        uint8_t *t = st + 14 * k;
                        uint8_t bIdx = t[8] & 7, aIdx = (t[8] >> 3) & 7;
                        if (nIn > 1 && nOut > 1) { memset(t, 0, 8); continue; }
                        int bad = 0;
                        if ((nIn > 1 && bIdx != maxIn) || (nOut > 1 && aIdx != maxOut)) bad = 1;
                        else if ((double)sInR > (ORC_MUTANT == 3 ? 0.25 : 0.2) ||
                                 (double)sOutR > (ORC_MUTANT == 3 ? 0.25 : 0.2))
                            bad = 1;
                        if (bad) {
                            uint64_t id = t_label(t);
                            m->parent[id] = ORC_BAD;
                            m->filtered[id] = 1;
                        }
                    }
                }
                curK = nk;
                curOfs = j;
                deg = 0;
                for (int l = 0; l < 7; l++) before[l] = after[l] = 0;
            }
            uint8_t a = st[14 * j + 8];
            before[a & 7]++;
            after[(a >> 3) & 7]++;
            deg++;
        }
    }
    if (scan >= 0 && scan < ORC_MAX_SCANS) {
        g_runs_eval[scan] = nev;
        g_runs_zero[scan] = nz;
    }
    /* 2500-2517 */
    for (uint64_t i = 0; i < m->N; i++)
        if (m->parent[i] != ORC_BAD && m->filtered[i]) m->parent[i] = ORC_BAD;
}

/* unionFindPar2 (metapart.c:2520-2735); the edge log and the race repair of
 * unionFind (2736-2800) have no effect without concurrent threads. */
static int mp_union(mp_t *m, const uint8_t *st, uint64_t numKmers, const uint64_t *lrc, const uint64_t *roff)
{
    if (numKmers == 0) return 0;
    /* every thread reads these before any thread writes the tuple array */
    uint64_t lastK = t_kmer(st + 14 * (numKmers - 1));
    uint64_t L0 = t_label(st);
    for (int seg = 0; seg < m->T; seg++) {
        uint64_t curK = lastK, minId = L0;
        for (uint64_t i = roff[seg]; i < roff[seg] + lrc[seg]; i++) {
            const uint8_t *t = st + 14 * i;
            uint64_t nk = t_kmer(t);
            uint64_t id = t_label(t);
            if (id >= m->N) return fail(-1, "label out of range");
            if (m->parent[id] == ORC_BAD) continue;
            if (nk == 0) continue;
            if (nk != curK) { curK = nk; minId = id; continue; }
            uint64_t pu = minId, pv = id;
            if (pu != pv) {
                if (pu >= m->N) return fail(-1, "label out of range");
                pu = find_root(m->parent, pu);
                pv = find_root(m->parent, pv);
            }
            if (pu != pv) {
                if (pu > pv) m->parent[pv] = pu;
                else m->parent[pu] = pv;
            }
        }
    }
    return 0;
}

/* printStats (metapart.c:3227-3308): compress and count sizes */
static void mp_compress(mp_t *m)
{
    for (uint64_t i = 0; i < m->N; i++) {
        uint64_t p = m->parent[i];
        if (p != ORC_BAD) {
            p = find_root(m->parent, i);
            m->parent[i] = p;
        }
        if (p == i) continue;
        if (p != ORC_BAD) { m->sizes[i] = 0; m->sizes[p]++; }
    }
}

/* unionFindPeReads (metapart.c:2801-2867) */
static void mp_pair_union(mp_t *m, uint32_t upto)
{
    for (uint64_t i = 0; i + 1 < m->N + 1 && i < m->N; i += 2) {
        uint64_t a = i, b = i + 1;
        if (m->parent[a] == ORC_BAD) m->parent[a] = a;
        if (m->parent[b] == ORC_BAD) m->parent[b] = b;
        uint64_t ra = find_root(m->parent, a), rb = find_root(m->parent, b);
        const int small_a = m->sizes[ra] <= upto, small_b = m->sizes[rb] <= upto;
        if (ORC_MUTANT == 1 ? (small_a && small_b) : (small_a || small_b)) {
            if (ra != rb) {
                if (ra > rb) { m->parent[rb] = ra; m->sizes[ra] += m->sizes[rb]; m->sizes[rb] = 0; }
                else { m->parent[ra] = rb; m->sizes[rb] += m->sizes[ra]; m->sizes[ra] = 0; }
            }
        }
    }
}

/* writeOutput3 (metapart.c:956-1066) */
static int mp_write(mp_t *m, const char *out_prefix, const uint64_t *W)
{
    int T = m->T, NP = 16;
    FILE **fo = (FILE **)calloc((size_t)NP * T, sizeof(FILE *));
    char name[4200];
    int rc = 0;
    for (int j = 0; j < NP * T; j++) {
        snprintf(name, sizeof(name), "%s/Rank0/FromTid%d_ToRank%d.paired.fastq", out_prefix, j % T, j / T);
        fo[j] = fopen(name, "wb");
        if (!fo[j]) { rc = fail(-1, "cannot create %s", name); goto out; }
    }
    int cpt = m->C / T;
    for (int t = 0; t < T && !rc; t++) {
        int interleave = 0;
        uint64_t ilStart = 0;
        for (int c = t * cpt; c < (t + 1) * cpt && !rc; c++) {
            uint64_t id = (uint64_t)m->fI[c].readId + (uint64_t)m->fI[c].readIdExtra * U32MAX;
            char *b;
            uint64_t len;
            if ((rc = mp_load_chunk(m, c, &b, &len, &interleave, &ilStart))) break;
            uint64_t ind = 0, r1, l1, r2, l2;
            int ret;
            while ((ret = next_full_read(b, len, &ind, interleave, ilStart, &r1, &l1, &r2, &l2)) != -1) {
                if (id >= m->N || (ret == 2 && id + 1 > m->N)) { rc = fail(-1, "read id past N"); break; }
                uint64_t pu = m->parent[id];
                int P = 0;
                while (P < NP && !(pu < W[P + 1])) P++;
                if (P >= NP || pu < W[P]) { rc = fail(-1, "partition lookup failed"); break; }
                if (ret == 2) {
                    uint64_t mate = (id % 2 == 0) ? id + 1 : id - 1;
                    uint64_t pv = m->parent[mate];
                    int Q = 0;
                    while (Q < NP && !(pv < W[Q + 1])) Q++;
                    if (Q >= NP || pv < W[Q]) { rc = fail(-1, "partition lookup failed"); break; }
                    fwrite(b + r1, 1, l1, fo[P * T + t]);
                    fwrite(b + r2, 1, l2, fo[P * T + t]);
                    if (pu != pv) {
                        fwrite(b + r1, 1, l1, fo[Q * T + t]);
                        fwrite(b + r2, 1, l2, fo[Q * T + t]);
                    }
                } else {
                    fwrite(b + r1, 1, l1, fo[P * T + t]);
                }
                id += ret == 2 ? 2 : 1;
            }
            free(b);
        }
    }
out:
    for (int j = 0; j < NP * T; j++)
        if (fo[j]) fclose(fo[j]);
    free(fo);
    return rc;
}

static void mp_release(mp_t *m)
{
    free(m->fsize); free(m->fI); free(m->H); free(m->scan); free(m->thread);
    free(m->tkc); free(m->sendThreadSizes); free(m->parent); free(m->filtered); free(m->sizes);
}

int orc_metapart(const char *out_prefix, uint64_t buf_size, int K, int S, int T, int nPE, int nSE, int nIL,
                 const char *const *files, int nfiles, int min3, uint64_t **parent_out, uint64_t *num_reads_out)
{
    (void)nIL;
    mp_t mm, *m = &mm;
    memset(m, 0, sizeof(*m));
    memset(g_runs_eval, 0, sizeof(g_runs_eval));
    memset(g_runs_zero, 0, sizeof(g_runs_zero));
    if (K <= 10 || K > 32) return fail(-1, "K must be in [11, 32]");
    if (S <= 0 || T <= 0 || T > 256) return fail(-1, "bad S/T");
    if (nPE > 0) return fail(-1, "two-file paired-end input is not supported");
    m->files = files; m->nfiles = nfiles; m->K = K; m->S = S; m->T = T; m->min3 = min3; m->bufsz = buf_size;
    m->numPEFiles = 2 * nPE; m->numSEFiles = nSE;
    m->fsize = (uint64_t *)calloc((size_t)nfiles, sizeof(uint64_t));
    for (int i = 0; i < nfiles; i++)
        if (file_size(files[i], &m->fsize[i])) { mp_release(m); return fail(-1, "cannot stat %s", files[i]); }
    char path[4096];
    snprintf(path, sizeof(path), min3 ? "%s.KmerIdx_3min%d" : "%s.KmerIdx%d", files[0], K);
    uint64_t isz;
    if (file_size(path, &isz) || isz < 2 * REC_SIZE || isz % REC_SIZE) { mp_release(m); return fail(-1, "bad %s", path); }
    m->C = (int)(isz / REC_SIZE) - 1;
    if (m->C % T) { mp_release(m); return fail(-1, "numThreads must divide the chunk count"); } /* 3657 */
    m->fI = (orc_fileindex *)malloc((size_t)(m->C + 1) * REC_SIZE);
    FILE *fp = fopen(path, "rb");
    size_t got = fread(m->fI, REC_SIZE, (size_t)m->C + 1, fp);
    fclose(fp);
    if (got != (size_t)m->C + 1) { mp_release(m); return fail(-1, "short %s", path); }
    m->N = (uint64_t)m->fI[m->C].readIdExtra * U32MAX + m->fI[m->C].readId;
    if (m->N == 0) { mp_release(m); return fail(-1, "no reads"); }
    if (m->N & 1) { mp_release(m); return fail(-1, "odd read count (the pair step reads id N)"); }
    m->H = (uint64_t *)malloc(sizeof(uint64_t) * (ORC_NBUCKETS + 1));
    snprintf(path, sizeof(path), min3 ? "%s.kmerHist_3min%d" : "%s.kmerHist%d", files[0], K);
    fp = fopen(path, "rb");
    if (!fp || fread(m->H, sizeof(uint64_t), ORC_NBUCKETS + 1, fp) != ORC_NBUCKETS + 1) {
        if (fp) fclose(fp);
        mp_release(m);
        return fail(-1, "bad %s", path);
    }
    fclose(fp);
    if (m->H[ORC_NBUCKETS] == 0) { mp_release(m); return fail(-1, "no k-mers"); }
    m->scan = (uint64_t *)calloc((size_t)S + 1, sizeof(uint64_t));
    m->thread = (uint64_t *)calloc((size_t)S * T + 1, sizeof(uint64_t));
    int rc = orc_plan_ranges(m->H, S, T, m->scan, m->thread);
    if (rc) { mp_release(m); return rc; }
    for (int i = 0; i < S * T; i++)
        if (!(m->thread[i] < m->thread[i + 1])) { mp_release(m); return fail(-3, "degenerate range plan (empty or unordered thread range)"); }
    /* 3793-3843 */
    m->sendThreadSizes = (uint64_t *)calloc((size_t)S * T, sizeof(uint64_t));
    m->tkc = (uint64_t *)calloc((size_t)T * ORC_NBUCKETS, sizeof(uint64_t));
    int cpt = m->C / T;
    for (int t = 0; t < T; t++)
        for (int c = t * cpt; c < (t + 1) * cpt; c++) {
            uint32_t k = 0;
            for (int s = 0; s < S; s++)
                while (k < m->scan[s + 1]) { m->sendThreadSizes[(size_t)s * T + t] += m->fI[c].kmerFreqCount[k]; k++; }
            for (uint32_t b = 0; b < ORC_NBUCKETS; b++) m->tkc[(size_t)t * ORC_NBUCKETS + b] += m->fI[c].kmerFreqCount[b];
        }
    m->filtered = (uint8_t *)calloc(m->N, 1);
    m->sizes = (uint32_t *)malloc(sizeof(uint32_t) * m->N);
    uint64_t *per_thread = (uint64_t *)calloc((size_t)T, sizeof(uint64_t));
    uint64_t *lrc = (uint64_t *)calloc((size_t)T + 1, sizeof(uint64_t));
    uint64_t *roff = (uint64_t *)calloc((size_t)T, sizeof(uint64_t));
    for (int s = 0; s < S && !rc; s++) { /* 3923-3974 */
        uint8_t *tuples = NULL, *sorted = NULL;
        uint64_t n = 0;
        if ((rc = mp_enumerate(m, s, &tuples, &n, per_thread))) break;
        if (m->parent == NULL) {
            m->parent = (uint64_t *)malloc(sizeof(uint64_t) * m->N);
            for (uint64_t i = 0; i < m->N; i++) { m->parent[i] = i; m->sizes[i] = 1; }
        }
        rc = mp_sort(m, s, tuples, n, per_thread, &sorted, lrc, roff);
        free(tuples);
        if (rc) break;
        mp_filter(m, s, sorted, lrc, roff);
        rc = mp_union(m, sorted, n, lrc, roff);
        if (!rc) mp_compress(m);
        if (s == g_dump_scan) {
            free(g_dump_tuples); free(g_dump_parent);
            g_dump_tuples = sorted; sorted = NULL; g_dump_n = n;
            g_dump_parent = (uint64_t *)malloc(sizeof(uint64_t) * m->N);
            memcpy(g_dump_parent, m->parent, sizeof(uint64_t) * m->N);
            g_dump_nreads = m->N;
        }
        free(sorted);
    }
    free(per_thread); free(lrc); free(roff);
    if (rc) { mp_release(m); return rc; }
    /* 4010-4056 */
    for (uint64_t i = 0; i < m->N; i++) m->sizes[i] = 1;
    mp_compress(m);
    mp_pair_union(m, 1000);
    for (uint64_t i = 0; i < m->N; i++) m->sizes[i] = 1;
    mp_compress(m);
    uint64_t W[17];
    memset(W, 0, sizeof(W));
    uint64_t X = m->N / 16, curTotal = 0, curInd = 0;
    W[16] = m->N;
    for (int i = 1; i < 16; i++) {
        while (curInd < m->N && curTotal < (uint64_t)i * X) {
            if (m->parent[curInd] == curInd) curTotal += m->sizes[curInd];
            curInd++;
        }
        W[i] = curInd;
    }
    rc = mp_write(m, out_prefix, W);
    if (!rc && parent_out) {
        *parent_out = m->parent;
        m->parent = NULL;
        if (num_reads_out) *num_reads_out = m->N;
    }
    mp_release(m);
    return rc;
}
