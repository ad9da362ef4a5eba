/* Seeded synthetic metagenome FASTQ generator.
 *
 * Produces single-end (or interleaved paired) FASTQ in the record layout the
 * partitioning tools expect: "@r<i>\n<seq>\n+\n<qual>\n", quality constant 'I',
 * an even number of records.  Every random draw is a pure function of
 * (seed, genome index) or (seed, read index), so the output does not depend on
 * the number of OpenMP threads used to make it.
 */
#ifndef MP_SYNTH_H
#define MP_SYNTH_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t seed;
    int      n_genomes;
    uint64_t genome_len_min, genome_len_max;  /* uniform, inclusive           */
    double   gc_min, gc_max;                  /* per-genome GC fraction       */
    double   abund_sigma;                     /* LogNormal(0, sigma); 0 = all equal */
    int      n_repeat_families;               /* shared repeats               */
    int      repeat_len;                      /* bases per repeat             */
    int      repeat_copies;                   /* genomes each family is planted in */
    uint64_t n_reads;                         /* rounded up to even           */
    int      read_len_min, read_len_max;      /* per-read length, uniform     */
    double   err_rate;                        /* per-base substitution prob.  */
    int      interleaved;                     /* 1: mates "/1","/2" of one fragment */
    int      insert_len;                      /* fragment length for pairs    */
    double   n_rate;                          /* per-base prob. of an 'N'     */
} mp_synth_cfg;

/* Fills *cfg with one of the named workloads (1..5 as in the configuration
 * table of the benchmark; 0 = tiny test pool).  Returns 0 on success. */
int mp_synth_preset(int which, mp_synth_cfg *cfg);

/* Generates the FASTQ into a malloc'd buffer (*out, *out_len).  Returns 0 on
 * success, <0 on bad configuration or allocation failure. */
int mp_synth_generate(const mp_synth_cfg *cfg, char **out, size_t *out_len);

/* Writes the FASTQ to a file. */
int mp_synth_write(const mp_synth_cfg *cfg, const char *path);

void mp_synth_free(char *buf);

#ifdef __cplusplus
}
#endif
#endif
