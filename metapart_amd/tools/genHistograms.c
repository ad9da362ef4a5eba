/* genHistograms_metapart[min] kmerSize numPEFiles numSEFiles numInterleavedFiles f1 [f2 ...]
 * Reads <f1>.idx, counts canonical k-mers per m-mer bucket per chunk on the
 * GPU and writes <f1>.KmerIdx<K> and <f1>.kmerHist<K> (…_3min<K> when built
 * with -DMIN3), as genHistograms.c does. */
#include "mp_tool_common.h"

#ifdef MIN3
#define VARIANT "_3min"
#define IS_MIN3 1
#else
#define VARIANT ""
#define IS_MIN3 0
#endif

int main(int argc, char **argv)
{
    if (argc < 6) {
        printf("Usage: %s kmerSize numPEFiles numSEFiles numInterleavedFiles a.1.fastq a.2.fastq b.1.fastq "
               "b.2.fastq ...\n", argv[0]);
        return 1;
    }
    const int K = atoi(argv[1]);
    const int nPE = atoi(argv[2]), nSE = atoi(argv[3]), nIL = atoi(argv[4]);
    (void)nSE; (void)nIL;
    const char *const *files = (const char *const *)&argv[5];
    const int nfiles = argc - 5;
    if (nPE > 0) { fprintf(stderr, "two-file paired-end input is not supported\n"); return 1; }
    char path[4096];
    snprintf(path, sizeof(path), "%s.idx", files[0]);
    int nrec = 0;
    mp_fileindex_hdr *hdr = NULL;
    if (read_index(path, &nrec, &hdr, NULL)) { fprintf(stderr, "cannot read %s\n", path); return 2; }
    const int C = nrec - 1;
    uint64_t *fstart = (uint64_t *)malloc(sizeof(uint64_t) * (nfiles + 1)), total = 0;
    char *fq = load_files(files, nfiles, fstart, &total);
    if (!fq) { fprintf(stderr, "cannot load FASTQ\n"); return 2; }
    uint64_t *cofs = (uint64_t *)malloc(sizeof(uint64_t) * (C + 1));
    for (int c = 0; c <= C; c++) cofs[c] = fstart[hdr[c].fileNo] + hdr[c].offset;
    mp_cfg cfg;
    memset(&cfg, 0, sizeof(cfg));
    cfg.k = K;
    cfg.min3 = IS_MIN3;
    mp_ctx *ctx = NULL;
    int rc = mp_ctx_create(&cfg, &ctx);
    if (rc) { fprintf(stderr, "mp_ctx_create: error %d\n", rc); return 2; }
    uint32_t *counts = (uint32_t *)malloc((size_t)C * MP_FILEINDEX_COUNTS * 4);
    uint64_t *reads = (uint64_t *)calloc((size_t)C + 1, sizeof(uint64_t)), kmers = 0;
    double t0 = now_s();
    rc = mp_count_chunks(ctx, fq, total, cofs, C, counts, reads, &kmers);
    double t1 = now_s();
    if (rc) { fprintf(stderr, "genHistograms: %s\n", mp_last_error(ctx)); return 2; }
    /* read-id bases: chunk c starts at the number of reads before it (genHistograms.c:280-291) */
    uint64_t rc_acc = 0;
    for (int i = 1; i <= C; i++) {
        rc_acc += reads[i - 1];
        hdr[i].readId = (uint32_t)(rc_acc % U32MAX);
        if (rc_acc / U32MAX >= 256) { fprintf(stderr, "too many reads\n"); return 2; }
        hdr[i].readIdExtra = (uint8_t)(rc_acc / U32MAX);
    }
    snprintf(path, sizeof(path), "%s.KmerIdx" VARIANT "%d", files[0], K);
    FILE *o = fopen(path, "wb");
    if (!o) { fprintf(stderr, "cannot create %s\n", path); return 2; }
    for (int c = 0; c <= C; c++)
        if (mp_fileindex_write(o, &hdr[c], c < C ? counts + (size_t)c * MP_FILEINDEX_COUNTS : NULL)) {
            fprintf(stderr, "write error\n");
            return 2;
        }
    fclose(o);
    /* exclusive prefix over buckets (genHistograms.c:321-343) */
    uint64_t *H = (uint64_t *)calloc(MP_NBUCKETS + 1, sizeof(uint64_t));
    for (int c = 0; c < C; c++) {
        const uint32_t *cc = counts + (size_t)c * MP_FILEINDEX_COUNTS;
        for (uint32_t b = 0; b < MP_NBUCKETS; b++) H[b + 1] += cc[b];
    }
    for (uint32_t b = 1; b <= MP_NBUCKETS; b++) H[b] += H[b - 1];
    if (H[MP_NBUCKETS] != kmers) { fprintf(stderr, "histogram total mismatch\n"); return 2; }
    snprintf(path, sizeof(path), "%s.kmerHist" VARIANT "%d", files[0], K);
    o = fopen(path, "wb");
    if (!o || fwrite(H, 8, MP_NBUCKETS + 1, o) != MP_NBUCKETS + 1) { fprintf(stderr, "cannot write %s\n", path); return 2; }
    fclose(o);
    printf("[0]TOTALKMERS %lu, TOTALREADS %lu\n", (unsigned long)kmers, (unsigned long)rc_acc);
    printf("[0] count %lf sec\n", t1 - t0);
    mp_ctx_destroy(ctx);
    free(H); free(counts); free(reads); free(cofs); free(fq); free(fstart); free(hdr);
    return 0;
}
