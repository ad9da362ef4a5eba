// Host-side parts of the C ABI that need no GPU: the FASTQ chunk indexer and
// the range planner.  Both must reproduce the reference's integer walks
// exactly because their outputs decide read ids, scan ranges and thread
// segments; the loop structure therefore follows the cited reference code.
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/metapart_amd.h"
#include "mp_fileindex.h"

namespace {

// genFastqIdx.c:28-53: byte position per*chunk -> (file, offset); a file pair
// of two-file paired-end input counts twice and the offset is halved.
bool locate(uint64_t pos, const std::vector<uint64_t> &fsize, int numPE, int32_t *file, uint64_t *off)
{
    int cur = 0;
    int factor = cur < numPE ? 2 : 1;
    while (cur < (int)fsize.size() && pos >= fsize[cur] * (uint64_t)factor) {
        pos -= fsize[cur] * (uint64_t)factor;
        cur += cur < numPE ? 2 : 1;
        factor = cur < numPE ? 2 : 1;
    }
    if (cur >= (int)fsize.size()) return false;
    *file = cur;
    *off = pos / (uint64_t)factor;
    return true;
}


using Probe = std::function<size_t(int file, uint64_t off, uint8_t *dst, size_t n)>;
using Sink = std::function<int(const mp_fileindex_hdr &)>;

// genFastqIdx.c:142-376 for single-end and interleaved files
int index_core(int num_chunks, int numPE, int numSE, const std::vector<uint64_t> &fsize, const Probe &probe,
               const Sink &sink)
{
    const int nfiles = (int)fsize.size();
    uint64_t total = 0;
    for (uint64_t v : fsize) total += v;
    const uint64_t per = total / (uint64_t)num_chunks;
    mp_fileindex_hdr hdr;
    memset(&hdr, 0, sizeof(hdr));
    std::vector<uint8_t> probe_buf(5000 + 64);
    int rc = 0, interleaved = 0;
    int64_t maxRead = 300;  // genFastqIdx.c:144, tightened after every chunk (303)
    for (int i = 0; i < num_chunks && !rc; i++) {
        if (!locate(per * (uint64_t)i, fsize, numPE, &hdr.fileNo, &hdr.offset)) { rc = MP_EINVAL; break; }
        if (hdr.fileNo >= numPE + numSE) interleaved = 1;
        if (fsize[hdr.fileNo] - hdr.offset < (uint64_t)maxRead) {  // near the end of a file (219-225)
            if (hdr.fileNo + 1 == nfiles) break;
            hdr.fileNo++;
            hdr.offset = 0;
        }
        if (i == 0) {
            hdr.offset2 = 0;
            hdr.readId = 0;
            rc = sink(hdr);
            continue;
        }
        std::fill(probe_buf.begin(), probe_buf.end(), 0);
        const size_t nread = probe(hdr.fileNo, hdr.offset, probe_buf.data(), 3000);
        if (!(nread > (uint64_t)maxRead)) { rc = MP_EINVAL; break; }
        const uint8_t *b = probe_buf.data();
        const size_t lim = 4990;
        bool found = false;
        size_t j = 0;
        int64_t k = 0;
        // first "\n+" line, skip it and the quality line: the next byte starts a
        // record; interleaved input moves on by one record when the two names
        // (up to ' ' or '/') differ, so a chunk starts at a pair (257-302)
        for (j = 0; j < nread && !rc; j++) {
            if (b[j] == '\n' && b[j + 1] == '+') {
                found = true;
                j += 2;
                while (j < lim && b[j] != '\n') j++;
                j++;
            }
            k = 0;
            if (found && j < lim && b[j] == '\n') {
                j++;
                k = 1;
                if (interleaved) {
                    size_t q = j;
                    while (q < lim && b[q] != '/' && b[q] != ' ') q++;
                    const std::string n1((const char *)b + j, q - j);
                    for (int l = 0; l < 4; l++) {
                        while (q < lim && b[q] != '\n') q++;
                        q++;
                    }
                    const size_t j2 = q;
                    while (q < lim && b[q] != ' ' && b[q] != '/') q++;
                    if (q >= lim) { rc = MP_EINVAL; break; }
                    const std::string n2((const char *)b + j2, q - j2);
                    if (n1 != n2) j = j2;
                    k = (int64_t)q;
                }
                break;
            }
            if (j >= lim) { rc = MP_EINVAL; break; }
        }
        if (rc) break;
        if (2 * k + 30 < maxRead) maxRead = 2 * k + 30;
        if (!found) { rc = MP_EINVAL; break; }
        hdr.offset += j;
        hdr.offset2 = 0;
        hdr.readId = 0;
        rc = sink(hdr);
    }
    if (!rc) {  // sentinel: end of the last file (338-361)
        hdr.fileNo = nfiles - 1;
        hdr.offset = fsize[nfiles - 1];
        hdr.offset2 = 0;
        hdr.readId = 0;
        rc = sink(hdr);
    }
    return rc;
}

}  // namespace

extern "C" int mp_index_fastq(int num_chunks, int nPE, int nSE, int nIL, const char *const *files, int nfiles)
{
    (void)nIL;
    if (num_chunks <= 0 || nfiles <= 0 || !files) return MP_EINVAL;
    if (nPE > 0) return MP_EINVAL;  // two-file paired-end input is not supported
    std::vector<uint64_t> fsize(nfiles);
    for (int i = 0; i < nfiles; i++) {
        struct stat st;
        if (stat(files[i], &st) != 0) return MP_EIO;
        fsize[i] = (uint64_t)st.st_size;
    }
    std::string path = std::string(files[0]) + ".idx";
    FILE *out = fopen(path.c_str(), "wb");
    if (!out) return MP_EIO;
    Probe probe = [&](int file, uint64_t off, uint8_t *dst, size_t n) -> size_t {
        FILE *fp = fopen(files[file], "rb");
        if (!fp) return 0;
        fseek(fp, (long)off, SEEK_SET);
        const size_t got = fread(dst, 1, n, fp);
        fclose(fp);
        return got;
    };
    Sink sink = [&](const mp_fileindex_hdr &h) -> int { return mp_fileindex_write(out, &h, nullptr); };
    int rc = index_core(num_chunks, 0, nSE, fsize, probe, sink);
    if (fclose(out) != 0 && !rc) rc = MP_EIO;
    if (rc) remove(path.c_str());  // no partial index for a later step to read
    return rc;
}

extern "C" int mp_index_buffer(const char *buf, uint64_t nbytes, int num_chunks, int interleaved, uint64_t *chunk_ofs,
                               int *n_chunks)
{
    if (!buf || !chunk_ofs || !n_chunks || num_chunks <= 0) return MP_EINVAL;
    std::vector<uint64_t> fsize(1, nbytes);
    Probe probe = [&](int, uint64_t off, uint8_t *dst, size_t n) -> size_t {
        if (off >= nbytes) return 0;
        const size_t got = (size_t)std::min<uint64_t>(n, nbytes - off);
        memcpy(dst, buf + off, got);
        return got;
    };
    int count = 0;
    Sink sink = [&](const mp_fileindex_hdr &h) -> int {
        if (count > num_chunks) return MP_EINVAL;
        chunk_ofs[count++] = h.offset;
        return 0;
    };
    const int rc = index_core(num_chunks, 0, interleaved ? 0 : 1, fsize, probe, sink);
    if (rc) return rc;
    *n_chunks = count - 1;
    return 0;
}

// metapart.c:3720-3791 with one MPI rank: the process ranges equal the scan
// ranges (same thresholds), so only the scan and thread walks remain.  Each
// walk moves at most one boundary per bucket, and the thread walk copies the
// scan boundary at every multiple of T.
extern "C" int mp_plan_ranges(const uint64_t *H, int S, int T, uint64_t *scan, uint64_t *thread)
{
    if (!H || !scan || !thread || S <= 0 || T <= 0) return MP_EINVAL;
    const uint64_t tot = H[MP_NBUCKETS];
    uint64_t avgS = tot / (uint64_t)S + (tot % (uint64_t)S ? 1 : 0);
    memset(scan, 0, sizeof(uint64_t) * (size_t)(S + 1));
    int r = 1;
    uint32_t i;
    for (i = 0; i < MP_NBUCKETS; i++)
        if (H[i] >= avgS * (uint64_t)r) {
            if (r > S) return MP_EPLAN;
            scan[r++] = i;
        }
    if (r == S) scan[r] = i;
    const uint64_t avgT = avgS / (uint64_t)T;
    memset(thread, 0, sizeof(uint64_t) * (size_t)(S * T + 1));
    r = 1;
    for (i = 0; i < MP_NBUCKETS; i++)
        if (H[i] >= (uint64_t)r * avgT) {
            if (r > S * T) return MP_EPLAN;
            thread[r] = (r % T == 0) ? scan[r / T] : i;
            r++;
        }
    if (r == S * T) thread[r] = i;
    for (int q = 0; q < S * T; q++)
        if (!(thread[q] < thread[q + 1])) return MP_EPLAN;
    return 0;
}
