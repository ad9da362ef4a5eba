// FASTQ record table and 2-bit sequence packing.
//
// The reference parses records byte by byte on the host for every scan
// (getNextRead, metapart.c:296-369; genHistograms.c:60-80).  Here the FASTQ
// bytes are uploaded once: newlines are counted per 16 KiB block, scanned, and
// every newline learns its global index j; for 4-line records, newline j ends
// line j%4 of record j/4.  From that each record gets its sequence offset and
// length, and the sequences are packed into a 2-bit stream (A0 C1 G2 T3,
// first base in the most significant bits, the reference's
// charToIntMap[(c&6)>>1]) plus a 1-bit stream marking the exact byte 'N',
// which is what the reference's lastNPos logic tests.
#include "mp_internal.h"
#include "mp_gather.h"

namespace mp {

constexpr int kParseBlock = 256;
constexpr int kParseBytesPerThread = 64;
constexpr uint64_t kParseBlockBytes = (uint64_t)kParseBlock * kParseBytesPerThread;  // 16 KiB

__device__ __forceinline__ int count_nl16(uint4 v)
{
    int c = 0;
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
        uint32_t x = w[i] ^ 0x0A0A0A0Au;  // zero bytes where the byte was '\n'
        uint32_t z = (x - 0x01010101u) & ~x & 0x80808080u;
        // exact per-byte zero test (the borrow trick can flag a 0x01 byte after a zero byte)
        if (z) {
#pragma unroll
            for (int b = 0; b < 4; b++) c += ((w[i] >> (8 * b)) & 0xFF) == '\n';
        }
    }
    return c;
}

__global__ __launch_bounds__(kParseBlock) void nl_count_kernel(const uint8_t *__restrict__ fq, uint64_t nbytes,
                                                               uint32_t *__restrict__ cnt)
{
    const uint64_t base = (uint64_t)blockIdx.x * kParseBlockBytes + (uint64_t)threadIdx.x * kParseBytesPerThread;
    int c = 0;
    if (base + kParseBytesPerThread <= nbytes) {
        const uint4 *p = (const uint4 *)(fq + base);
#pragma unroll
        for (int i = 0; i < 4; i++) c += count_nl16(p[i]);
    } else {
        for (uint64_t i = base; i < nbytes && i < base + kParseBytesPerThread; i++) c += fq[i] == '\n';
    }
    // block reduction
    for (int d = kWave / 2; d > 0; d >>= 1) c += __shfl_down(c, d, kWave);
    __shared__ int s[kParseBlock / kWave];
    if (__lane_id() == 0) s[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kParseBlock / kWave; w++) t += s[w];
        cnt[blockIdx.x] = (uint32_t)t;
    }
}

// Writes, for newline j at byte p: j%4==0 -> seq_off[j/4] = p+1;
// j%4==1 -> seq_end[j/4] = p; j%4==2 -> plus_end[j/4] = p; j%4==3 -> rec_off[j/4+1] = p+1.
// It also checks the byte after the sequence line ('+') and after a record
// ('@' of the next record): first_bad takes the lowest record that fails.
__global__ __launch_bounds__(kParseBlock) void nl_write_kernel(const uint8_t *__restrict__ fq, uint64_t nbytes,
                                                               const uint64_t *__restrict__ blk_off, uint64_t n_rec,
                                                               uint64_t *__restrict__ seq_off,
                                                               uint64_t *__restrict__ seq_end,
                                                               uint64_t *__restrict__ plus_end,
                                                               uint64_t *__restrict__ rec_off,
                                                               unsigned long long *__restrict__ first_bad)
{
    const uint64_t base = (uint64_t)blockIdx.x * kParseBlockBytes + (uint64_t)threadIdx.x * kParseBytesPerThread;
    uint32_t w[16];
    int n = 0;
    if (base + kParseBytesPerThread <= nbytes) {
        const uint4 *p = (const uint4 *)(fq + base);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            uint4 v = p[i];
            w[4 * i] = v.x; w[4 * i + 1] = v.y; w[4 * i + 2] = v.z; w[4 * i + 3] = v.w;
        }
        n = kParseBytesPerThread;
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++) w[i] = 0;
        for (uint64_t i = base; i < nbytes && i < base + kParseBytesPerThread; i++) n++;
#pragma unroll
        for (int i = 0; i < 64; i++)
            if (i < n) w[i >> 2] |= (uint32_t)fq[base + i] << (8 * (i & 3));
    }
    // bit i of nl: byte i is a newline (exact zero-byte test on each word xor '\n')
    uint64_t nl = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t x = w[i] ^ 0x0A0A0A0Au;
        const uint32_t y = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);  // 0x80 where the byte is 0
        const uint32_t nib = ((y >> 7) & 1u) | ((y >> 14) & 2u) | ((y >> 21) & 4u) | ((y >> 28) & 8u);
        nl |= (uint64_t)nib << (4 * i);
    }
    if (n < 64) nl &= (1ull << n) - 1;
    const int c = __popcll(nl);
    // exclusive prefix of c over the block
    const unsigned lane = __lane_id();
    int inc = c;
    for (int d = 1; d < kWave; d <<= 1) {
        int o = __shfl_up(inc, d, kWave);
        if (lane >= (unsigned)d) inc += o;
    }
    __shared__ int s[kParseBlock / kWave];
    if (lane == kWave - 1) s[threadIdx.x / kWave] = inc;
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < (int)(threadIdx.x / kWave); w++) woff += s[w];
    uint64_t j = blk_off[blockIdx.x] + (uint64_t)(woff + inc - c);
    for (uint64_t m = nl; m; m &= m - 1) {
        const int i = __builtin_ctzll(m);
        const uint64_t pos = base + (uint64_t)i;
        const uint64_t r = j >> 2;
        if (r < n_rec) {
            switch (j & 3) {
            case 0: seq_off[r] = pos + 1; break;
            case 1: seq_end[r] = pos; break;
            case 2: plus_end[r] = pos; break;
            default: rec_off[r + 1] = pos + 1; break;
            }
            if ((j & 3) == 1 || ((j & 3) == 3 && r + 1 < n_rec)) {
                uint32_t nb;  // the byte after the newline
                if (i + 1 < n) {
                    const int k = i + 1;
                    uint32_t wk = 0;
#pragma unroll
                    for (int t = 0; t < 16; t++) wk = t == (k >> 2) ? w[t] : wk;
                    nb = (wk >> (8 * (k & 3))) & 0xFF;
                } else {
                    nb = pos + 1 < nbytes ? fq[pos + 1] : 0u;
                }
                const uint32_t want = (j & 3) == 1 ? '+' : '@';
                if (nb != want) atomicMin(first_bad, (unsigned long long)((j & 3) == 1 ? r : r + 1));
            }
        }
        j++;
    }
}

// Per record: sequence length, k-mer position count, and the checks that the
// reference's parsers assume silently (4-line records, '@' header, '+' line,
// |qual| == |seq|, at most 299+K sequence bytes after the first k-mer).
__global__ void rec_finalize_kernel(const uint8_t *__restrict__ fq, uint64_t n_rec, int K,
                                    const uint64_t *__restrict__ seq_off, const uint64_t *__restrict__ seq_end,
                                    const uint64_t *__restrict__ plus_end, const uint64_t *__restrict__ rec_off,
                                    uint32_t *__restrict__ seq_len, uint32_t *__restrict__ npos,
                                    unsigned long long *__restrict__ first_bad)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rec) return;
    const uint64_t so = seq_off[r], se = seq_end[r], pe = plus_end[r], re = rec_off[r + 1], rs = rec_off[r];
    const uint64_t L = se - so;
    const uint64_t qual = re - 1 - (pe + 1);
    // the '+' line and the '@' of every later record were checked by nl_write_kernel
    bool ok = (r != 0 || fq[rs] == '@') && qual == L && L < 65536;
    if (ok && L > (uint64_t)K && L - K >= 300) ok = false;  // enumerateKmers.c:155
    if (!ok) atomicMin(first_bad, (unsigned long long)r);
    seq_len[r] = (uint32_t)L;
    npos[r] = L > (uint64_t)K ? (uint32_t)(L - K + 1) : 0u;
}

// Packing: the sequence bytes of the reads are gathered in base order (read
// r's bases at base_off[r]) and turned into the 2-bit and 'N' streams in the
// same pass, 16 bases per lane (mp_gather.h).
struct SeqSrc {
    const uint64_t *seq_off;
    __device__ uint64_t operator()(uint64_t r) const { return seq_off[r]; }
};

// chunk_ofs[c] -> index of the record starting there (or n_rec for the end);
// flags a chunk boundary that is not a record start.
__global__ void chunk_ids_kernel(const uint64_t *__restrict__ rec_off, uint64_t n_rec,
                                 const uint64_t *__restrict__ chunk_ofs, int C, uint64_t *__restrict__ ids,
                                 unsigned long long *__restrict__ first_bad)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > C) return;
    const uint64_t target = chunk_ofs[c];
    uint64_t lo = 0, hi = n_rec;  // first r with rec_off[r] >= target
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (rec_off[mid] < target) lo = mid + 1; else hi = mid;
    }
    ids[c] = lo;
    if (rec_off[lo] != target) atomicMin(first_bad, (unsigned long long)(1ull << 62) + (unsigned long long)c);
}

int parse_reads(mp_ctx *ctx, const uint64_t *chunk_ofs, const uint64_t *chunk_first_id, int C)
{
    const uint64_t nbytes = ctx->fq_bytes;
    const uint8_t *fq = ctx->d_fq;
    const int K = ctx->cfg.k;
    if (nbytes == 0) return ctx->fail(MP_EINVAL, "empty FASTQ");
    if (((uintptr_t)fq) & 15) return ctx->fail(MP_EINVAL, "FASTQ device buffer must be 16-byte aligned");
    const uint64_t nblk = (nbytes + kParseBlockBytes - 1) / kParseBlockBytes;
    if (!grid_fits(nblk, kParseBlock)) return ctx->fail(MP_EINVAL, "FASTQ of %llu bytes exceeds one dispatch",
                                                        (unsigned long long)nbytes);

    // newline counts and offsets (scratch space: wg_cnt / wg_off)
    MP_TRY(ensure(ctx, ctx->wg_cnt, nblk * sizeof(uint32_t)));
    MP_TRY(ensure(ctx, ctx->wg_off, (nblk + 1) * sizeof(uint64_t)));
    hipLaunchKernelGGL(nl_count_kernel, dim3((unsigned)nblk), dim3(kParseBlock), 0, ctx->stream, fq, nbytes,
                       (uint32_t *)ctx->wg_cnt.p);
    MP_HIP(hipGetLastError());
    MP_TRY(scan_u32(ctx, (const uint32_t *)ctx->wg_cnt.p, (uint64_t *)ctx->wg_off.p, nblk));
    uint64_t n_nl = 0;
    MP_HIP(hipMemcpyAsync(&n_nl, (uint64_t *)ctx->wg_off.p + nblk, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    if (n_nl == 0 || (n_nl & 3))
        return ctx->fail(MP_EINVAL, "FASTQ has %llu lines; records must be 4 lines ending in a newline",
                         (unsigned long long)n_nl);
    const uint64_t N = n_nl / 4;
    if (N >= 0xFFFFFFFFull) return ctx->fail(MP_EINVAL, "too many reads for 32-bit ids");
    ctx->N = N;  // an odd count loads and counts; mp_scan refuses it

    MP_TRY(ensure(ctx, ctx->rec_off, (N + 1) * sizeof(uint64_t)));
    MP_TRY(ensure(ctx, ctx->seq_off, N * sizeof(uint64_t)));
    MP_TRY(ensure(ctx, ctx->seq_len, N * sizeof(uint32_t)));
    MP_TRY(ensure(ctx, ctx->base_off, (N + 1) * sizeof(uint64_t)));
    MP_TRY(ensure(ctx, ctx->pos_off, (N + 1) * sizeof(uint64_t)));
    // temporaries in the sort buffers (sized for at least N u64 each)
    MP_TRY(ensure(ctx, ctx->keys[0], N * sizeof(uint64_t)));
    MP_TRY(ensure(ctx, ctx->keys[1], N * sizeof(uint64_t)));
    MP_TRY(ensure(ctx, ctx->lab[0], N * sizeof(uint32_t)));
    uint64_t *seq_end = (uint64_t *)ctx->keys[0].p, *plus_end = (uint64_t *)ctx->keys[1].p;
    uint32_t *npos = (uint32_t *)ctx->lab[0].p;
    uint64_t *rec_off = (uint64_t *)ctx->rec_off.p;
    MP_TRY(fill_u32(ctx, rec_off, 0, 2));
    MP_TRY(ensure(ctx, ctx->hist, 4096));
    unsigned long long *first_bad = (unsigned long long *)ctx->hist.p;
    MP_TRY(fill_u32(ctx, first_bad, 0xFFFFFFFFu, 2));
    hipLaunchKernelGGL(nl_write_kernel, dim3((unsigned)nblk), dim3(kParseBlock), 0, ctx->stream, fq, nbytes,
                       (const uint64_t *)ctx->wg_off.p, N, (uint64_t *)ctx->seq_off.p, seq_end, plus_end, rec_off,
                       first_bad);
    MP_HIP(hipGetLastError());
    hipLaunchKernelGGL(rec_finalize_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, ctx->stream, fq, N, K,
                       (const uint64_t *)ctx->seq_off.p, seq_end, plus_end, (const uint64_t *)rec_off,
                       (uint32_t *)ctx->seq_len.p, npos, first_bad);
    MP_HIP(hipGetLastError());
    unsigned long long bad = 0;
    MP_HIP(hipMemcpyAsync(&bad, first_bad, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream));
    uint64_t last_end = 0;
    MP_HIP(hipMemcpyAsync(&last_end, rec_off + N, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    if (bad != ~0ull)
        return ctx->fail(MP_EINVAL, "malformed FASTQ record %llu (needs '@' header, '+' line, |qual|==|seq|, "
                                    "|seq|-K < 300)", bad);
    if (last_end != nbytes) return ctx->fail(MP_EINVAL, "trailing bytes after the last FASTQ record");

    // chunk boundaries must be record starts; ids must agree with the caller's
    MP_TRY(ensure(ctx, ctx->scratch, 0));
    {
        uint64_t *d_cofs = nullptr, *d_ids = nullptr;
        MP_TRY(tmp_alloc(ctx, (void **)&d_cofs, sizeof(uint64_t) * 2 * (C + 1)));
        d_ids = d_cofs + (C + 1);
        MP_HIP(hipMemcpyAsync(d_cofs, chunk_ofs, sizeof(uint64_t) * (C + 1), hipMemcpyHostToDevice, ctx->stream));
        MP_TRY(fill_u32(ctx, first_bad, 0xFFFFFFFFu, 2));
        hipLaunchKernelGGL(chunk_ids_kernel, dim3((C + 256) / 256), dim3(256), 0, ctx->stream, (const uint64_t *)rec_off,
                           N, (const uint64_t *)d_cofs, C, d_ids, first_bad);
        std::vector<uint64_t> ids(C + 1);
        MP_HIP(hipMemcpyAsync(ids.data(), d_ids, sizeof(uint64_t) * (C + 1), hipMemcpyDeviceToHost, ctx->stream));
        MP_HIP(hipMemcpyAsync(&bad, first_bad, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream));
        MP_TRY(tmp_free(ctx, d_cofs));
        MP_HIP(hipStreamSynchronize(ctx->stream));
        if (bad != ~0ull) return ctx->fail(MP_EINVAL, "chunk %llu does not start at a FASTQ record", bad - (1ull << 62));
        if (ids[C] != N) return ctx->fail(MP_EINVAL, "chunks do not end at the last record");
        if (chunk_first_id) {
            for (int c = 0; c <= C; c++)
                if (ids[c] != chunk_first_id[c])
                    return ctx->fail(MP_EINVAL, "chunk %d: index says first read %llu, FASTQ has %llu", c,
                                     (unsigned long long)chunk_first_id[c], (unsigned long long)ids[c]);
        }
        ctx->C = C;
        ctx->chunk_ids.assign(ids.begin(), ids.end());
    }

    // base and k-mer position offsets
    MP_TRY(scan_u32(ctx, (const uint32_t *)ctx->seq_len.p, (uint64_t *)ctx->base_off.p, N));
    MP_TRY(scan_u32(ctx, npos, (uint64_t *)ctx->pos_off.p, N));
    uint64_t tb = 0;
    MP_HIP(hipMemcpyAsync(&tb, (uint64_t *)ctx->base_off.p + N, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    ctx->total_bases = tb;
    const uint64_t nthreads = (tb + 63) / 64 + 2;  // two zero words of padding after the stream
    MP_TRY(ensure(ctx, ctx->packed, nthreads * 2 * sizeof(uint64_t)));
    MP_TRY(ensure(ctx, ctx->nmask, nthreads * sizeof(uint64_t)));
    uint64_t *first = nullptr;
    const uint64_t nb = (tb + kGBlock - 1) / kGBlock;
    MP_TRY(tmp_alloc(ctx, (void **)&first, (nb + 1) * sizeof(uint64_t)));
    const int rc = launch_gather(ctx, fq, ctx->fq_bytes, (const uint64_t *)ctx->base_off.p, N, tb, nthreads * 64, first,
                                 SeqSrc{(const uint64_t *)ctx->seq_off.p},
                                 PackSink{(uint32_t *)ctx->packed.p, (uint16_t *)ctx->nmask.p});
    MP_TRY(tmp_free(ctx, first));
    MP_TRY(rc);
    MP_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // namespace mp
