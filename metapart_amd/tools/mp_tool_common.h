/* Shared helpers of the command-line tools: file loading and index I/O. */
#ifndef MP_TOOL_COMMON_H
#define MP_TOOL_COMMON_H
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>

#include "../../include/metapart_amd.h"
#include "../csrc/mp_fileindex.h"

#define U32MAX 4294967295ull

static inline double now_s(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + tv.tv_usec * 1e-6;
}

static inline int file_size(const char *p, uint64_t *sz)
{
    struct stat st;
    if (stat(p, &st) != 0) return -1;
    *sz = (uint64_t)st.st_size;
    return 0;
}

/* All files back to back in argv order; fstart[i] = offset of file i. */
static inline char *load_files(const char *const *files, int nfiles, uint64_t *fstart, uint64_t *total)
{
    uint64_t t = 0;
    for (int i = 0; i < nfiles; i++) {
        uint64_t s;
        if (file_size(files[i], &s)) { fprintf(stderr, "cannot stat %s\n", files[i]); return NULL; }
        fstart[i] = t;
        t += s;
    }
    fstart[nfiles] = t;
    char *buf = (char *)malloc(t + 64);
    if (!buf) return NULL;
    for (int i = 0; i < nfiles; i++) {
        FILE *fp = fopen(files[i], "rb");
        if (!fp) { free(buf); return NULL; }
        uint64_t want = fstart[i + 1] - fstart[i];
        if (fread(buf + fstart[i], 1, want, fp) != want) { fclose(fp); free(buf); return NULL; }
        fclose(fp);
    }
    *total = t;
    return buf;
}

/* Reads every record header (and optionally all counts) of an index file. */
static inline int read_index(const char *path, int *nrec_out, mp_fileindex_hdr **hdr, uint32_t **counts)
{
    uint64_t sz;
    if (file_size(path, &sz) || sz % MP_FILEINDEX_BYTES || sz < 2 * MP_FILEINDEX_BYTES) return -1;
    int n = (int)(sz / MP_FILEINDEX_BYTES);
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    *hdr = (mp_fileindex_hdr *)calloc((size_t)n, sizeof(mp_fileindex_hdr));
    if (counts) *counts = (uint32_t *)malloc((size_t)n * MP_FILEINDEX_COUNTS * 4);
    for (int i = 0; i < n; i++)
        if (mp_fileindex_read(fp, &(*hdr)[i], counts ? *counts + (size_t)i * MP_FILEINDEX_COUNTS : NULL)) {
            fclose(fp);
            return -1;
        }
    fclose(fp);
    *nrec_out = n;
    return 0;
}
#endif
