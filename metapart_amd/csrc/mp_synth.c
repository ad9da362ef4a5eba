/* Seeded synthetic metagenome FASTQ generator (see mp_synth.h). */
#include "mp_synth.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline uint64_t sm64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline uint64_t stream_seed(uint64_t seed, uint64_t tag, uint64_t idx)
{
    uint64_t s = seed * 0xD1B54A32D192ED03ull ^ (tag << 56) ^ idx;
    (void)sm64(&s);
    return sm64(&s);
}
static inline double unif(uint64_t *s) { return (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0); }
static inline uint64_t below(uint64_t *s, uint64_t n) { return n ? sm64(s) % n : 0; }
static inline double gauss(uint64_t *s)
{
    double u1 = unif(s), u2 = unif(s);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static const char BASES[4] = {'A', 'C', 'G', 'T'};
static inline char comp(char c)
{
    switch (c) { case 'A': return 'T'; case 'C': return 'G'; case 'G': return 'C'; case 'T': return 'A'; default: return c; }
}

int mp_synth_preset(int which, mp_synth_cfg *c)
{
    memset(c, 0, sizeof(*c));
    c->gc_min = c->gc_max = 0.5;
    c->insert_len = 300;
    switch (which) {
    case 0: /* tiny pool for unit tests */
        c->seed = 7; c->n_genomes = 4; c->genome_len_min = 3000; c->genome_len_max = 6000;
        c->abund_sigma = 1.0; c->n_reads = 2000; c->read_len_min = c->read_len_max = 100; c->err_rate = 0.005;
        return 0;
    case 1: /* 16 genomes 40-200 kb iid ACGT, LogNormal(0,1), 250k x 100 bp, 0.5 % */
        c->seed = 1; c->n_genomes = 16; c->genome_len_min = 40000; c->genome_len_max = 200000;
        c->abund_sigma = 1.0; c->n_reads = 250000; c->read_len_min = c->read_len_max = 100; c->err_rate = 0.005;
        return 0;
    case 2: /* 200 genomes 1-5 Mb, GC U(0.3,0.7), 10 x 5 kb repeats in 2 genomes each, 20M x 150 bp, 1 % */
        c->seed = 2; c->n_genomes = 200; c->genome_len_min = 1000000; c->genome_len_max = 5000000;
        c->gc_min = 0.3; c->gc_max = 0.7; c->abund_sigma = 0.0;
        c->n_repeat_families = 10; c->repeat_len = 5000; c->repeat_copies = 2;
        c->n_reads = 20000000; c->read_len_min = c->read_len_max = 150; c->err_rate = 0.01;
        return 0;
    case 3: /* 1000 genomes 0.5-8 Mb, LogNormal(0,1.5), 100M x 150 bp, 1 % */
        c->seed = 3; c->n_genomes = 1000; c->genome_len_min = 500000; c->genome_len_max = 8000000;
        c->abund_sigma = 1.5; c->n_reads = 100000000; c->read_len_min = c->read_len_max = 150; c->err_rate = 0.01;
        return 0;
    default:
        return -1;
    }
}

typedef struct {
    char    *seq;     /* all genomes back to back */
    uint64_t *start;  /* n_genomes + 1 */
    double  *cumw;    /* cumulative sampling weight, n_genomes */
} pool_t;

static int make_pool(const mp_synth_cfg *c, pool_t *p)
{
    int G = c->n_genomes;
    p->start = (uint64_t *)malloc(sizeof(uint64_t) * (G + 1));
    p->cumw = (double *)malloc(sizeof(double) * G);
    double *gc = (double *)malloc(sizeof(double) * G);
    if (!p->start || !p->cumw || !gc) return -2;
    p->start[0] = 0;
    double acc = 0;
    for (int g = 0; g < G; g++) {
        uint64_t s = stream_seed(c->seed, 1, (uint64_t)g);
        uint64_t span = c->genome_len_max - c->genome_len_min + 1;
        uint64_t len = c->genome_len_min + below(&s, span);
        gc[g] = c->gc_min + (c->gc_max - c->gc_min) * unif(&s);
        double ab = c->abund_sigma > 0 ? exp(c->abund_sigma * gauss(&s)) : 1.0;
        p->start[g + 1] = p->start[g] + len;
        acc += ab * (double)len;
        p->cumw[g] = acc;
    }
    p->seq = (char *)malloc(p->start[G] + 1);
    if (!p->seq) { free(gc); return -2; }
#pragma omp parallel for schedule(dynamic, 1)
    for (int g = 0; g < G; g++) {
        const double thr = gc[g];
        /* 64 blocks per genome so big genomes still spread over threads */
        uint64_t a = p->start[g], b = p->start[g + 1];
        for (uint64_t blk = 0; blk * 65536 < b - a; blk++) {
            uint64_t s = stream_seed(c->seed, 2, ((uint64_t)g << 32) | blk);
            uint64_t lo = a + blk * 65536, hi = lo + 65536 < b ? lo + 65536 : b;
            for (uint64_t i = lo; i < hi; i++) {
                uint64_t r = sm64(&s);
                double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
                int isgc = u < thr;
                int bit = (int)(r & 1);
                p->seq[i] = isgc ? (bit ? 'G' : 'C') : (bit ? 'T' : 'A');
            }
        }
    }
    free(gc);
    /* shared repeats: family f is planted in repeat_copies distinct genomes */
    if (c->n_repeat_families > 0 && c->repeat_len > 0) {
        char *rep = (char *)malloc((size_t)c->repeat_len);
        if (!rep) return -2;
        for (int f = 0; f < c->n_repeat_families; f++) {
            uint64_t s = stream_seed(c->seed, 3, (uint64_t)f);
            for (int i = 0; i < c->repeat_len; i++) rep[i] = BASES[sm64(&s) & 3];
            int placed = 0, tries = 0;
            int used[64];
            while (placed < c->repeat_copies && placed < 64 && tries++ < 10000) {
                int g = (int)below(&s, (uint64_t)G);
                int dup = 0;
                for (int q = 0; q < placed; q++) dup |= used[q] == g;
                uint64_t len = p->start[g + 1] - p->start[g];
                if (dup || len < (uint64_t)c->repeat_len) continue;
                uint64_t pos = p->start[g] + below(&s, len - c->repeat_len + 1);
                memcpy(p->seq + pos, rep, (size_t)c->repeat_len);
                used[placed++] = g;
            }
        }
        free(rep);
    }
    return 0;
}

static int ndigits(uint64_t v)
{
    int d = 1;
    while (v >= 10) { v /= 10; d++; }
    return d;
}

/* length of record i (header "@r<id>" [+ "/1"|"/2"]) */
static inline uint64_t rec_len(const mp_synth_cfg *c, uint64_t i, int L)
{
    uint64_t id = c->interleaved ? i / 2 : i;
    uint64_t hdr = 2 + (uint64_t)ndigits(id) + (c->interleaved ? 2 : 0);
    return hdr + 1 + (uint64_t)L + 1 + 2 + (uint64_t)L + 1;
}

static inline int read_len_of(const mp_synth_cfg *c, uint64_t i)
{
    if (c->read_len_max <= c->read_len_min) return c->read_len_min;
    uint64_t frag = c->interleaved ? i / 2 : i;
    uint64_t s = stream_seed(c->seed, 4, frag);
    return c->read_len_min + (int)below(&s, (uint64_t)(c->read_len_max - c->read_len_min + 1));
}

static void emit_record(const mp_synth_cfg *c, const pool_t *p, uint64_t i, int L, char *dst)
{
    uint64_t frag = c->interleaved ? i / 2 : i;
    int mate = c->interleaved ? (int)(i & 1) : 0;
    uint64_t s = stream_seed(c->seed, 5, frag);
    double u = unif(&s) * p->cumw[c->n_genomes - 1];
    int lo = 0, hi = c->n_genomes - 1;
    while (lo < hi) { int mid = (lo + hi) / 2; if (p->cumw[mid] > u) hi = mid; else lo = mid + 1; }
    int g = lo;
    uint64_t glen = p->start[g + 1] - p->start[g];
    int strand = (int)(sm64(&s) & 1);
    int span = c->interleaved ? (c->insert_len > L ? c->insert_len : L) : L;
    if ((uint64_t)span > glen) span = (int)glen;
    uint64_t fpos = glen > (uint64_t)span ? below(&s, glen - (uint64_t)span + 1) : 0;
    const char *gs = p->seq + p->start[g];
    /* mate 0 reads the fragment start on `strand`; mate 1 the other end, opposite strand */
    int rev = mate ? !strand : strand;
    uint64_t off;
    if (!c->interleaved) off = fpos;
    else off = (mate == strand) ? fpos : fpos + (uint64_t)span - (uint64_t)L;
    if (L > (int)glen) L = (int)glen;
    uint64_t es = stream_seed(c->seed, 6, i);
    char *q = dst;
    uint64_t id = c->interleaved ? frag : i;
    q += sprintf(q, "@r%llu", (unsigned long long)id);
    if (c->interleaved) { *q++ = '/'; *q++ = mate ? '2' : '1'; }
    *q++ = '\n';
    const uint64_t thr = (uint64_t)(c->err_rate * 18446744073709551616.0);
    const uint64_t nthr = (uint64_t)(c->n_rate * 18446744073709551616.0);
    for (int j = 0; j < L; j++) {
        char b = rev ? comp(gs[off + (uint64_t)(L - 1 - j)]) : gs[off + (uint64_t)j];
        uint64_t r = sm64(&es);
        if (c->err_rate > 0 && r < thr) {
            int k = 0;
            while (k < 4 && BASES[k] != b) k++;
            int nb = (int)((k + 1 + (sm64(&es) % 3)) & 3);
            b = BASES[nb];
        }
        if (c->n_rate > 0 && sm64(&es) < nthr) b = 'N';
        q[j] = b;
    }
    q += L;
    *q++ = '\n'; *q++ = '+'; *q++ = '\n';
    memset(q, 'I', (size_t)L);
    q += L;
    *q++ = '\n';
}

int mp_synth_generate(const mp_synth_cfg *c, char **out, size_t *out_len)
{
    if (!c || !out || !out_len || c->n_genomes <= 0 || c->read_len_min <= 0 ||
        c->read_len_max < c->read_len_min || c->genome_len_min == 0 ||
        c->genome_len_max < c->genome_len_min)
        return -1;
    uint64_t N = c->n_reads + (c->n_reads & 1);
    pool_t p;
    memset(&p, 0, sizeof(p));
    int rc = make_pool(c, &p);
    if (rc) { free(p.seq); free(p.start); free(p.cumw); return rc; }
    uint64_t *ofs = (uint64_t *)malloc(sizeof(uint64_t) * (N + 1));
    if (!ofs) { free(p.seq); free(p.start); free(p.cumw); return -2; }
    ofs[0] = 0;
    for (uint64_t i = 0; i < N; i++) ofs[i + 1] = ofs[i] + rec_len(c, i, read_len_of(c, i));
    char *buf = (char *)malloc(ofs[N] + 64);
    if (!buf) { free(ofs); free(p.seq); free(p.start); free(p.cumw); return -2; }
#pragma omp parallel for schedule(static, 4096)
    for (uint64_t i = 0; i < N; i++) emit_record(c, &p, i, read_len_of(c, i), buf + ofs[i]);
    *out = buf;
    *out_len = ofs[N];
    free(ofs); free(p.seq); free(p.start); free(p.cumw);
    return 0;
}

int mp_synth_write(const mp_synth_cfg *c, const char *path)
{
    char *buf = NULL;
    size_t len = 0;
    int rc = mp_synth_generate(c, &buf, &len);
    if (rc) return rc;
    FILE *f = fopen(path, "wb");
    if (!f) { free(buf); return -3; }
    size_t w = fwrite(buf, 1, len, f);
    fclose(f);
    free(buf);
    return w == len ? 0 : -3;
}

void mp_synth_free(char *buf) { free(buf); }
