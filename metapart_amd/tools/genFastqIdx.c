/* genFastqIdx numChunks numPEFiles numSEFiles numInterleavedFiles f1 [f2 ...]
 * Same command line and <f1>.idx output as the reference tool (genFastqIdx.c). */
#include "mp_tool_common.h"

int main(int argc, char **argv)
{
    if (argc < 6) {
        printf("Usage: %s numPartitions numPEFiles numSEFiles numInterleavedFiles a.1.fastq a.2.fastq b.1.fastq "
               "b.2.fastq ...\n", argv[0]);
        return 1;
    }
    int C = atoi(argv[1]);
    if (C <= 0) { fprintf(stderr, "numPartitions must be > 0\n"); return 1; }
    int rc = mp_index_fastq(C, atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), (const char *const *)&argv[5], argc - 5);
    if (rc) {
        fprintf(stderr, "genFastqIdx: error %d (%s)\n", rc, rc == MP_EINVAL ? "unsupported or malformed input" : "i/o");
        return 2;
    }
    return 0;
}
