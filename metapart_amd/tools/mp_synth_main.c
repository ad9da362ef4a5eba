/* mp_synth: write a seeded synthetic FASTQ (workload presets 0-3 plus overrides). */
#include "../csrc/mp_synth.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void usage(const char *a)
{
    fprintf(stderr,
            "Usage: %s [--preset P] [--seed S] [--reads N] [--genomes G] [--glen-min A] [--glen-max B]\n"
            "          [--len L | --len-min A --len-max B] [--err E] [--sigma S] [--gc-min A --gc-max B]\n"
            "          [--interleaved] [--nrate R] [--repeats F] -o out.fastq\n", a);
}

int main(int argc, char **argv)
{
    mp_synth_cfg c;
    const char *out = NULL;
    int preset = 0;
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
        else if (!strcmp(a, "--repeats")) { c.n_repeat_families = atoi(v); if (!c.repeat_len) c.repeat_len = 5000; if (!c.repeat_copies) c.repeat_copies = 2; }
        else { usage(argv[0]); return 1; }
        i++;
    }
    if (!out) { usage(argv[0]); return 1; }
    int rc = mp_synth_write(&c, out);
    if (rc) fprintf(stderr, "mp_synth: error %d\n", rc);
    return rc ? 2 : 0;
}
