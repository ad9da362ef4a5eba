/* On-disk chunk record shared by <fastq>.idx and <fastq>.KmerIdx<K>[_3min].
 *
 * Byte layout of the reference's fileIndex struct on x86-64
 * (genFastqIdx.c:11-18), 4,194,344 bytes per record:
 *   0  u64 offset      8 u64 offset2     16 u8 readIdExtra   17..19 padding
 *   20 u32 readId     24 i32 fileNo      28 u32 kmerFreqCount[2^20 + 2]
 *   4,194,340..343 padding
 * The reference leaves the padding (and readIdExtra of record 0) as
 * uninitialised stack bytes; this build always writes zeros there.
 */
#ifndef MP_FILEINDEX_H
#define MP_FILEINDEX_H
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#define MP_FILEINDEX_BYTES 4194344ull
#define MP_FILEINDEX_COUNTS (1048576u + 2u)

typedef struct {
    uint64_t offset;
    uint64_t offset2;
    uint8_t readIdExtra;
    uint8_t pad[3];
    uint32_t readId;
    int32_t fileNo;
} mp_fileindex_hdr;

/* counts: MP_FILEINDEX_COUNTS values, or NULL for zeros.  Returns 0 or -6. */
static inline int mp_fileindex_write(FILE *fp, const mp_fileindex_hdr *h, const uint32_t *counts)
{
    uint8_t head[28];
    static const uint8_t zeros[65536] = {0};
    memset(head, 0, sizeof(head));
    memcpy(head + 0, &h->offset, 8);
    memcpy(head + 8, &h->offset2, 8);
    head[16] = h->readIdExtra;
    memcpy(head + 20, &h->readId, 4);
    memcpy(head + 24, &h->fileNo, 4);
    if (fwrite(head, 1, 28, fp) != 28) return -6;
    if (counts) {
        if (fwrite(counts, 4, MP_FILEINDEX_COUNTS, fp) != MP_FILEINDEX_COUNTS) return -6;
    } else {
        uint64_t left = 4ull * MP_FILEINDEX_COUNTS;
        while (left) {
            size_t n = left < sizeof(zeros) ? (size_t)left : sizeof(zeros);
            if (fwrite(zeros, 1, n, fp) != n) return -6;
            left -= n;
        }
    }
    if (fwrite(zeros, 1, 4, fp) != 4) return -6;
    return 0;
}

/* Reads one record; counts may be NULL to skip them.  Returns 0 or -6. */
static inline int mp_fileindex_read(FILE *fp, mp_fileindex_hdr *h, uint32_t *counts)
{
    uint8_t head[28];
    if (fread(head, 1, 28, fp) != 28) return -6;
    memset(h, 0, sizeof(*h));
    memcpy(&h->offset, head + 0, 8);
    memcpy(&h->offset2, head + 8, 8);
    h->readIdExtra = head[16];
    memcpy(&h->readId, head + 20, 4);
    memcpy(&h->fileNo, head + 24, 4);
    if (counts) {
        if (fread(counts, 4, MP_FILEINDEX_COUNTS, fp) != MP_FILEINDEX_COUNTS) return -6;
        uint8_t tail[4];
        if (fread(tail, 1, 4, fp) != 4) return -6;
    } else if (fseek(fp, (long)(MP_FILEINDEX_BYTES - 28), SEEK_CUR) != 0)
        return -6;
    return 0;
}
#endif
