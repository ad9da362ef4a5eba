/* viewIdx file.idx: one line per record (viewIdx.c output format). */
#include "mp_tool_common.h"

int main(int argc, char **argv)
{
    if (argc != 2) { fprintf(stderr, "%s fqIdxFile\n", argv[0]); return 1; }
    int n = 0;
    mp_fileindex_hdr *h = NULL;
    if (read_index(argv[1], &n, &h, NULL)) { fprintf(stderr, "cannot read %s\n", argv[1]); return 2; }
    for (int i = 0; i < n; i++)
        printf("%d %d %lu,%lu %u-%u\n", i, h[i].fileNo, (unsigned long)h[i].offset, (unsigned long)h[i].offset2,
               (unsigned)h[i].readIdExtra, h[i].readId);
    printf("TotalKmers %lu\n", 0ul);
    free(h);
    return 0;
}
