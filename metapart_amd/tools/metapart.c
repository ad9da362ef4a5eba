/* metapart[min] -o OUT fastqBufferSize kmerSize numScans numThreads numPEFiles numSEFiles numInterleavedFiles f1 ...
 *
 * Same command line, inputs (<f1>.KmerIdx<K>, <f1>.kmerHist<K>, or the _3min
 * pair when built with -DMIN3) and outputs (OUT/Rank0/FromTid<t>_ToRank<p>.paired.fastq)
 * as the reference partitioner (metapart.c:3544-4150) run as one MPI rank.
 * numThreads no longer sets a CPU thread count: it fixes the T segments whose
 * boundaries the reference's results depend on. */
#include <getopt.h>

#include "mp_tool_common.h"

#ifdef MIN3
#define VARIANT "_3min"
#define IS_MIN3 1
#else
#define VARIANT ""
#define IS_MIN3 0
#endif

static int die(mp_ctx *ctx, const char *what, int rc)
{
    fprintf(stderr, "metapart: %s failed (%d): %s\n", what, rc, ctx ? mp_last_error(ctx) : "");
    return 2;
}

int main(int argc, char **argv)
{
    char out_prefix[4096] = ".";
    static struct option opts[] = {{"out-prefix", required_argument, 0, 'o'}, {0, 0, 0, 0}};
    int c;
    while ((c = getopt_long(argc, argv, "o:", opts, NULL)) != -1)
        if (c == 'o') snprintf(out_prefix, sizeof(out_prefix), "%s", optarg);
    if (argc < 11 || argc - optind < 8) {
        printf("Usage: %s -o <output-prefix> fastqBufferSize kmerSize numScans numThreads numPEFiles numSEFiles "
               "numInterleavedFiles a.fastq b.fastq ...\n", argv[0]);
        return 1;
    }
    char **a = &argv[optind - 1];
    const uint64_t bufsz = strtoull(a[1], NULL, 10);
    const int K = atoi(a[2]), S = atoi(a[3]), T = atoi(a[4]);
    const int nPE = atoi(a[5]), nSE = atoi(a[6]), nIL = atoi(a[7]);
    const char *const *files = (const char *const *)&a[8];
    const int nfiles = argc - (optind + 7);
    if (S <= 0 || T <= 0 || nfiles <= 0) { fprintf(stderr, "bad numScans/numThreads/files\n"); return 1; }
    if (nPE > 0) { fprintf(stderr, "two-file paired-end input is not supported\n"); return 1; }
    const int numSE = nSE;
    (void)nIL;
    double t_start = now_s();

    /* KmerIdx and kmerHist (metapart.c:3641-3718) */
    char path[4096];
    snprintf(path, sizeof(path), "%s.KmerIdx" VARIANT "%d", files[0], K);
    int nrec = 0;
    mp_fileindex_hdr *hdr = NULL;
    uint32_t *counts = NULL;
    if (read_index(path, &nrec, &hdr, &counts)) { fprintf(stderr, "cannot read %s\n", path); return 2; }
    const int C = nrec - 1;
    if (C % T) { fprintf(stderr, "numThreads must divide the chunk count %d\n", C); return 2; }
    const uint64_t N = (uint64_t)hdr[C].readIdExtra * U32MAX + hdr[C].readId;
    uint64_t *H = (uint64_t *)malloc(sizeof(uint64_t) * (MP_NBUCKETS + 1));
    snprintf(path, sizeof(path), "%s.kmerHist" VARIANT "%d", files[0], K);
    FILE *fp = fopen(path, "rb");
    if (!fp || fread(H, 8, MP_NBUCKETS + 1, fp) != MP_NBUCKETS + 1) { fprintf(stderr, "cannot read %s\n", path); return 2; }
    fclose(fp);
    if (H[MP_NBUCKETS] == 0) { fprintf(stderr, "no k-mers\n"); return 2; }

    /* range plan and per-scan tuple counts (metapart.c:3720-3829) */
    uint64_t *scan = (uint64_t *)calloc((size_t)S + 1, 8), *thr = (uint64_t *)calloc((size_t)S * T + 1, 8);
    int rc = mp_plan_ranges(H, S, T, scan, thr);
    if (rc) { fprintf(stderr, "metapart: degenerate range plan for S=%d T=%d (%d)\n", S, T, rc); return 2; }
    uint64_t *expect = (uint64_t *)calloc((size_t)S, 8);
    for (int ch = 0; ch < C; ch++) {
        const uint32_t *cc = counts + (size_t)ch * MP_FILEINDEX_COUNTS;
        uint64_t k = 0;
        for (int s = 0; s < S; s++)
            while (k < scan[s + 1]) expect[s] += cc[k++];
    }
    free(counts);

    /* FASTQ bytes, chunk offsets and first read ids */
    uint64_t *fstart = (uint64_t *)malloc(sizeof(uint64_t) * (nfiles + 1)), total = 0;
    char *fq = load_files(files, nfiles, fstart, &total);
    if (!fq) { fprintf(stderr, "cannot load FASTQ\n"); return 2; }
    uint64_t *cofs = (uint64_t *)malloc(sizeof(uint64_t) * (C + 1));
    uint64_t *cid = (uint64_t *)malloc(sizeof(uint64_t) * (C + 1));
    for (int ch = 0; ch <= C; ch++) {
        cofs[ch] = fstart[hdr[ch].fileNo] + hdr[ch].offset;
        cid[ch] = (uint64_t)hdr[ch].readIdExtra * U32MAX + hdr[ch].readId;
    }
    cid[C] = N;
    for (int ch = 0; ch < C; ch++)
        if (!(bufsz > cofs[ch + 1] - cofs[ch])) { fprintf(stderr, "chunk %d larger than fastqBufferSize\n", ch); return 2; }
    /* interleaved start per chunk as loadChunk leaves it: sticky per thread (metapart.c:401-410) */
    int64_t *il = (int64_t *)malloc(sizeof(int64_t) * C);
    for (int t = 0; t < T; t++) {
        int interleave = 0;
        uint64_t ilStart = 0;
        for (int ch = t * (C / T); ch < (t + 1) * (C / T); ch++) {
            int f = hdr[ch].fileNo;
            uint64_t o = hdr[ch].offset, bo = 0;
            while (f < hdr[ch + 1].fileNo || (f == hdr[ch + 1].fileNo && o < hdr[ch + 1].offset)) {
                if (f >= numSE && !interleave) { interleave = 1; ilStart = bo; }
                uint64_t end = f < hdr[ch + 1].fileNo ? fstart[f + 1] - fstart[f] : hdr[ch + 1].offset;
                bo += end - o;
                f++;
                o = 0;
            }
            il[ch] = interleave ? (int64_t)ilStart : -1;
        }
    }

    mp_cfg cfg;
    memset(&cfg, 0, sizeof(cfg));
    cfg.k = K;
    cfg.min3 = IS_MIN3;
    mp_ctx *ctx = NULL;
    if ((rc = mp_ctx_create(&cfg, &ctx))) return die(NULL, "mp_ctx_create", rc);
    double t0 = now_s();
    if ((rc = mp_load_reads(ctx, fq, total, cofs, cid, C))) return die(ctx, "mp_load_reads", rc);
    double t1 = now_s();
    double gen = 0, sort = 0, filt = 0, uf = 0;
    for (int s = 0; s < S; s++) {
        mp_scan_stats st;
        if ((rc = mp_scan(ctx, s, scan[s], scan[s + 1], &thr[(size_t)s * T], T, expect[s], &st)))
            return die(ctx, "mp_scan", rc);
        gen += st.ms_enumerate / 1e3;
        sort += st.ms_sort / 1e3;
        filt += st.ms_filter / 1e3;
        uf += st.ms_union / 1e3;
        printf("[0] scan %d: %lu tuples, %d sort passes, %lu BAD reads, %lu zeroed runs\n", s, (unsigned long)st.tuples,
               st.sort_passes, (unsigned long)st.bad_labels, (unsigned long)st.zeroed_runs);
    }
    uint64_t W[17];
    mp_finish_stats fs;
    double t2 = now_s();
    if ((rc = mp_finish(ctx, W, &fs))) return die(ctx, "mp_finish", rc);
    double t3 = now_s();
    uint64_t outb = 0;
    if ((rc = mp_write_partitions(ctx, out_prefix, T, W, il, &outb))) return die(ctx, "mp_write_partitions", rc);
    double t4 = now_s();
    for (int p = 0; p <= 16; p++) printf("%lu\t", (unsigned long)W[p]);
    printf("\n");
    /* the reference's summary line (metapart.c:4100), device stage times */
    printf("[0] FQRead %lf sec(MaxKmerGen %lf sec), CommTime %lf sec, SortTime %lf sec(%lf + %lf), FilterTime %lf sec, "
           "unionFind %lf sec (%lf + %lf), mergeComps %lf sec, peUnionTime %lf sec, WriteFq %lf sec(%lf sec)\n",
           t1 - t0, gen, 0.0, sort, 0.0, sort, filt, uf, uf, 0.0, 0.0, t3 - t2, t4 - t3, 0.0);
    printf("[0] reads %lu, wall %lf sec, %.1f reads/s\n", (unsigned long)N, t4 - t_start, N / (t4 - t_start));
    mp_ctx_destroy(ctx);
    free(fq); free(fstart); free(cofs); free(cid); free(il); free(H); free(scan); free(thr); free(expect); free(hdr);
    return 0;
}
