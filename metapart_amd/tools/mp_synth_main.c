/* mp_synth: write a seeded synthetic FASTQ (workload presets 0-3 plus overrides). */
#include "../csrc/mp_synth.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void usage(const char *a)
{
    fprintf(stderr,
            "Usage: %s [--preset P] [--seed S] [--reads N] [--genomes G] [--glen-min A] [--glen-max B]\n"
            "          [--len L | --len-min A --len-max B] [--err E] [--sigma S] [--gc-min A --gc-max B]\n"
            "          [--interleaved] [--nrate R] [--repeats F] [--mix M] -o out.fastq\n", a);
}

/* --mix M: after writing, each sequence byte independently becomes lower case
 * with probability M, or one of the other IUPAC letters / a lower-case 'n'
 * with probability M/10 (a seeded stream of its own, so the file is otherwise
 * the same as without the option).  The reference reads any byte: its 2-bit
 * code takes (c & 6) >> 1, the adj byte takes the raw neighbours, and only an
 * upper-case 'N' stops a k-mer. */
static uint64_t mix_next(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);  /* splitmix64 */
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static int mix_file(const char *path, double m, uint64_t seed)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *b = (char *)malloc(n > 0 ? (size_t)n : 1);
    if (!b || fread(b, 1, (size_t)n, f) != (size_t)n) { fclose(f); free(b); return -1; }
    fclose(f);
    static const char other[] = "RYKMSWBDHVn";
    uint64_t st = seed ^ 0x6D69786D69786D69ull;
    const uint64_t lo = (uint64_t)(m * 18446744073709551615.0), oth = lo / 10;
    long line = 0;
    for (long i = 0; i < n; i++) {
        if (b[i] == '\n') { line++; continue; }
        if ((line & 3) != 1) continue;
        const uint64_t r = mix_next(&st);
        if (r < oth) b[i] = other[mix_next(&st) % (sizeof(other) - 1)];
        else if (r < oth + lo && b[i] >= 'A' && b[i] <= 'Z') b[i] = (char)(b[i] + 32);
    }
    f = fopen(path, "wb");
    const int ok = f && fwrite(b, 1, (size_t)n, f) == (size_t)n;
    if (f) fclose(f);
    free(b);
    return ok ? 0 : -1;
}

int main(int argc, char **argv)
{
    mp_synth_cfg c;
    const char *out = NULL;
    int preset = 0;
    double mix = 0;
    for (int i = 1; i < argc; i++)
        if (!strcmp(argv[i], "--preset") && i + 1 < argc) preset = atoi(argv[++i]);
    if (mp_synth_preset(preset, &c)) { usage(argv[0]); return 1; }
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        const char *v = i + 1 < argc ? argv[i + 1] : NULL;
        if (!strcmp(a, "--preset")) { i++; continue; }
        if (!strcmp(a, "--interleaved")) { c.interleaved = 1; continue; }
        if (!v) { usage(argv[0]); return 1; }
        if (!strcmp(a, "-o")) out = v;
        else if (!strcmp(a, "--seed")) c.seed = strtoull(v, NULL, 10);
        else if (!strcmp(a, "--reads")) c.n_reads = strtoull(v, NULL, 10);
        else if (!strcmp(a, "--genomes")) c.n_genomes = atoi(v);
        else if (!strcmp(a, "--glen-min")) c.genome_len_min = strtoull(v, NULL, 10);
        else if (!strcmp(a, "--glen-max")) c.genome_len_max = strtoull(v, NULL, 10);
        else if (!strcmp(a, "--len")) c.read_len_min = c.read_len_max = atoi(v);
        else if (!strcmp(a, "--len-min")) c.read_len_min = atoi(v);
        else if (!strcmp(a, "--len-max")) c.read_len_max = atoi(v);
        else if (!strcmp(a, "--err")) c.err_rate = atof(v);
        else if (!strcmp(a, "--sigma")) c.abund_sigma = atof(v);
        else if (!strcmp(a, "--gc-min")) c.gc_min = atof(v);
        else if (!strcmp(a, "--gc-max")) c.gc_max = atof(v);
        else if (!strcmp(a, "--nrate")) c.n_rate = atof(v);
        else if (!strcmp(a, "--mix")) mix = atof(v);
        else if (!strcmp(a, "--repeats")) { c.n_repeat_families = atoi(v); if (!c.repeat_len) c.repeat_len = 5000; if (!c.repeat_copies) c.repeat_copies = 2; }
        else { usage(argv[0]); return 1; }
        i++;
    }
    if (!out) { usage(argv[0]); return 1; }
    int rc = mp_synth_write(&c, out);
    if (!rc && mix > 0) rc = mix_file(out, mix, c.seed) ? 3 : 0;
    if (rc) fprintf(stderr, "mp_synth: error %d\n", rc);
    return rc ? 2 : 0;
}
