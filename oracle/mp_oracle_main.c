/* mp_oracle: command-line front end of the CPU oracle (test infrastructure).
 *   mp_oracle genFastqIdx C nPE nSE nIL f...
 *   mp_oracle genHistograms[min] K nPE nSE nIL f...
 *   mp_oracle metapart[min] -o OUT bufSz K S T nPE nSE nIL f...
 * Same argument vectors and output files as the reference tools. */
#include "mp_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv)
{
    if (argc < 2) { fprintf(stderr, "usage: %s tool args...\n", argv[0]); return 1; }
    const char *tool = argv[1];
    int rc;
    if (!strcmp(tool, "genFastqIdx")) {
        if (argc < 7) return 1;
        rc = orc_gen_fastq_idx(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]),
                               (const char *const *)&argv[6], argc - 6);
    } else if (!strcmp(tool, "genHistograms") || !strcmp(tool, "genHistogramsmin")) {
        if (argc < 7) return 1;
        rc = orc_gen_histograms(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]),
                                (const char *const *)&argv[6], argc - 6, !strcmp(tool, "genHistogramsmin"));
    } else if (!strcmp(tool, "metapart") || !strcmp(tool, "metapartmin")) {
        if (argc < 12 || strcmp(argv[2], "-o")) return 1;
        rc = orc_metapart(argv[3], strtoull(argv[4], NULL, 10), atoi(argv[5]), atoi(argv[6]), atoi(argv[7]),
                          atoi(argv[8]), atoi(argv[9]), atoi(argv[10]), (const char *const *)&argv[11], argc - 11,
                          !strcmp(tool, "metapartmin"), NULL, NULL);
    } else {
        fprintf(stderr, "unknown tool %s\n", tool);
        return 1;
    }
    if (rc) fprintf(stderr, "mp_oracle %s: error %d: %s\n", tool, rc, orc_last_error());
    return rc ? 2 : 0;
}
