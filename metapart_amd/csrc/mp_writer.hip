// Partitioned FASTQ writer (writeOutput3, metapart.c:956-1066).
//
// The reference re-reads every chunk on T threads and fwrite()s each record
// (single-end) or record pair (interleaved) to FromTid<t>_ToRank<P(root)>,
// a pair again to P(root of mate) when the two roots differ.  Here every
// output item (record or pair, copy 0 or 1) gets a file id f = P*T + t on the
// device, items are stably partitioned by f with the radix passes of the
// tuple sort (so each file keeps chunk and record order), byte offsets come
// from a scan over item lengths in that order, and one wave per item copies
// the record bytes into a single output buffer holding the 16*T files back to
// back.  The host then writes each file's slice.  A batch of samples (one
// chunk each) gets 16 files per sample, planned over each sample's own ids.
#include <sys/stat.h>

#include <cerrno>
#include <cstring>
#include <string>
#include <vector>

#include "mp_internal.h"
#include "mp_gather.h"

namespace mp {

constexpr uint32_t kNoFile = 0xFFFFu;

struct WParams {
    const uint64_t *rec_off;
    const uint32_t *parent;
    const uint64_t *chunk_ids;  // C+1
    const uint64_t *pair_from;  // per chunk: first record that starts pairs (or chunk end)
    const uint64_t *W;          // 17
    const uint64_t *Ws;         // batch of samples (one chunk each): 17 per sample, else null
    uint64_t N;
    int C, T, cpt;
    uint64_t *file;   // 2N item keys
    uint32_t *slot;   // 2N item payloads (item index)
    uint8_t *zero;    // 2N
    uint32_t *len;    // 2N item byte lengths
};

__device__ __forceinline__ int part_of(const uint64_t *W, uint64_t x)
{
    int p = 0;  // first p with x < W[p+1] (metapart.c:1011-1012)
    while (p < MP_NPART - 1 && !(x < W[p + 1])) p++;
    return p;
}

// first record of chunk c whose end offset (within the chunk) reaches il_start:
// records before it are single, records from it on pair up (getNextFullRead,
// metapart.c:211-294: a second record is read when the first ends at or after
// interleaveStart).
__global__ void pair_from_kernel(const uint64_t *__restrict__ rec_off, const uint64_t *__restrict__ chunk_ids,
                                 const uint64_t *__restrict__ chunk_ofs, const int64_t *__restrict__ il, int C,
                                 uint64_t *__restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    uint64_t lo = chunk_ids[c], hi = chunk_ids[c + 1];
    if (il[c] < 0) {
        out[c] = hi;
        return;
    }
    const uint64_t target = chunk_ofs[c] + (uint64_t)il[c];
    while (lo < hi) {  // first r with rec_off[r+1] >= target
        const uint64_t mid = (lo + hi) >> 1;
        if (rec_off[mid + 1] < target) lo = mid + 1; else hi = mid;
    }
    out[c] = lo;
}

__global__ void items_kernel(WParams P)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= P.N) return;
    int lo = 0, hi = P.C - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (P.chunk_ids[mid] <= r) lo = mid; else hi = mid - 1;
    }
    const int c = lo;
    const int t = c / P.cpt;
    const uint64_t pf = P.pair_from[c];
    // a sample of a batch has its own plan over its own read ids and files
    const uint64_t *W = P.Ws ? P.Ws + (size_t)c * (MP_NPART + 1) : P.W;
    const uint64_t id0 = P.Ws ? P.chunk_ids[c] : 0;
    const uint32_t fbase = P.Ws ? (uint32_t)c * MP_NPART * P.T : 0;
    uint32_t f0 = kNoFile, f1 = kNoFile, l0 = 0;
    const uint64_t pu = P.parent[r];
    if (r < pf) {
        f0 = fbase + (uint32_t)(part_of(W, pu - id0) * P.T + t);
        l0 = (uint32_t)(P.rec_off[r + 1] - P.rec_off[r]);
    } else if (((r - pf) & 1) == 0) {
        const uint64_t mate = (r & 1) == 0 ? r + 1 : r - 1;  // metapart.c:1013-1016
        const uint64_t pv = P.parent[mate];
        f0 = fbase + (uint32_t)(part_of(W, pu - id0) * P.T + t);
        if (pu != pv) f1 = fbase + (uint32_t)(part_of(W, pv - id0) * P.T + t);
        l0 = (uint32_t)(P.rec_off[r + 2] - P.rec_off[r]);
    }
    P.file[2 * r] = f0;
    P.file[2 * r + 1] = f1;
    P.slot[2 * r] = (uint32_t)(2 * r);
    P.slot[2 * r + 1] = (uint32_t)(2 * r + 1);
    P.zero[2 * r] = 0;
    P.zero[2 * r + 1] = 0;
    P.len[2 * r] = f0 == kNoFile ? 0 : l0;
    P.len[2 * r + 1] = f1 == kNoFile ? 0 : l0;
}

__global__ void gather_len_kernel(const uint32_t *__restrict__ slot, const uint32_t *__restrict__ len, uint64_t n,
                                  uint32_t *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = len[slot[i]];
}

// source of item i (in file order): its record, or the first record of its pair
struct RecSrc {
    const uint64_t *rec_off;
    const uint32_t *slot;
    __device__ uint64_t operator()(uint64_t i) const { return rec_off[slot[i] >> 1]; }
};

// first item of every file (and the end), and the byte offset there
__global__ void file_bounds_kernel(const uint64_t *__restrict__ file, uint64_t n, int F, const uint64_t *__restrict__ off,
                                   uint64_t *__restrict__ pos, uint64_t *__restrict__ byte)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f > F) return;
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (file[mid] < (uint64_t)f) lo = mid + 1; else hi = mid;
    }
    pos[f] = lo;
    byte[f] = off[lo];
}

static int make_dirs(const std::string &path)
{
    std::string cur;
    for (size_t i = 0; i <= path.size(); i++) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty() && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return -1;
        }
        if (i < path.size()) cur.push_back(path[i]);
    }
    return 0;
}

int write_partitions(mp_ctx *ctx, const char *out_prefix, int T, const uint64_t *W, const int64_t *il_start,
                     uint64_t *bytes_out)
{
    const uint64_t N = ctx->N;
    const int C = ctx->C;
    const bool batch = ctx->batch;
    const int F = MP_NPART * T * (batch ? C : 1);
    if (F >= (int)kNoFile) return ctx->fail(MP_EINVAL, "too many output files");
    // writer items are numbered 2 r and 2 r + 1 in 32 bits
    if (2 * N > (uint64_t)UINT32_MAX)
        return ctx->fail(MP_EINVAL, "the writer takes at most 2,147,483,647 reads (%llu loaded)",
                         (unsigned long long)N);
    if (batch && (T != 1 || ctx->sample_W.size() != (size_t)C * (MP_NPART + 1)))
        return ctx->fail(MP_ESTATE, "batch output needs T = 1 and per-sample plans");
    std::vector<int64_t> il(C, -1);
    if (il_start)
        for (int c = 0; c < C; c++) il[c] = il_start[c];
    // small device inputs
    const size_t nW = batch ? (size_t)C * (MP_NPART + 1) : (size_t)(MP_NPART + 1);
    const size_t nsmall = (size_t)(C + 1) * 2 + C + nW + (size_t)C + 2 * (size_t)(F + 1);
    uint64_t *d_small = nullptr;
    MP_HIP(hipMallocAsync((void **)&d_small, nsmall * sizeof(uint64_t), ctx->stream));
    uint64_t *d_ids = d_small, *d_cofs = d_small + (C + 1), *d_pf = d_small + 2 * (C + 1);
    uint64_t *d_W = d_pf + C;
    int64_t *d_il = (int64_t *)(d_W + nW);
    uint64_t *d_fpos = (uint64_t *)(d_il + C);
    uint64_t *d_fbyte = d_fpos + (F + 1);
    MP_HIP(hipMemcpyAsync(d_ids, ctx->chunk_ids.data(), sizeof(uint64_t) * (C + 1), hipMemcpyHostToDevice, ctx->stream));
    MP_HIP(hipMemcpyAsync(d_cofs, ctx->chunk_ofs.data(), sizeof(uint64_t) * (C + 1), hipMemcpyHostToDevice, ctx->stream));
    MP_HIP(hipMemcpyAsync(d_W, batch ? ctx->sample_W.data() : W, sizeof(uint64_t) * nW, hipMemcpyHostToDevice,
                          ctx->stream));
    MP_HIP(hipMemcpyAsync(d_il, il.data(), sizeof(int64_t) * C, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(pair_from_kernel, dim3((C + 255) / 256), dim3(256), 0, ctx->stream,
                       (const uint64_t *)ctx->rec_off.p, (const uint64_t *)d_ids, (const uint64_t *)d_cofs,
                       (const int64_t *)d_il, C, d_pf);
    MP_HIP(hipGetLastError());
    std::vector<uint64_t> pf(C);
    MP_HIP(hipMemcpyAsync(pf.data(), d_pf, sizeof(uint64_t) * C, hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < C; c++)
        if ((ctx->chunk_ids[c + 1] - pf[c]) & 1) {
            MP_HIP(hipFreeAsync(d_small, ctx->stream));
            return ctx->fail(MP_EINVAL, "chunk %d: interleaved part has an odd number of records", c);
        }
    const uint64_t n_items = 2 * N;
    for (int b = 0; b < 2; b++) {
        MP_TRY(ensure(ctx, ctx->items[b], n_items * sizeof(uint64_t)));
        MP_TRY(ensure(ctx, ctx->ilens[b], n_items * sizeof(uint32_t)));
        MP_TRY(ensure(ctx, ctx->iaux[b], n_items + 16));
    }
    MP_TRY(ensure(ctx, ctx->wg_cnt, n_items * sizeof(uint32_t)));
    MP_TRY(ensure(ctx, ctx->wg_off, (n_items + 1) * sizeof(uint64_t)));
    uint32_t *len = (uint32_t *)ctx->wg_cnt.p;
    WParams P;
    P.rec_off = (const uint64_t *)ctx->rec_off.p;
    P.parent = (const uint32_t *)ctx->parent.p;
    P.chunk_ids = d_ids;
    P.pair_from = d_pf;
    P.W = d_W;
    P.Ws = batch ? d_W : nullptr;
    P.N = N;
    P.C = C;
    P.T = T;
    P.cpt = C / T;
    P.file = (uint64_t *)ctx->items[0].p;
    P.slot = (uint32_t *)ctx->ilens[0].p;
    P.zero = (uint8_t *)ctx->iaux[0].p;
    P.len = len;
    hipEventRecord(ctx->ev[0], ctx->stream);
    hipLaunchKernelGGL(items_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, ctx->stream, P);
    MP_HIP(hipGetLastError());
    int cur = 0, passes = 0;
    MP_TRY(radix_sort_generic(ctx, ctx->items, ctx->ilens, ctx->iaux, &cur, n_items, 0x6 /* key bytes 0,1 */, &passes,
                                  nullptr, 0, nullptr));
    const uint64_t *file = (const uint64_t *)ctx->items[cur].p;
    const uint32_t *slot = (const uint32_t *)ctx->ilens[cur].p;
    // lengths in sorted order (reuse the other ilens buffer), offsets by scan
    uint32_t *slen = (uint32_t *)ctx->ilens[1 - cur].p;
    hipLaunchKernelGGL(gather_len_kernel, dim3((unsigned)((n_items + 255) / 256)), dim3(256), 0, ctx->stream, slot,
                       (const uint32_t *)len, n_items, slen);
    MP_HIP(hipGetLastError());
    uint64_t *off = (uint64_t *)ctx->wg_off.p;
    MP_TRY(scan_u32(ctx, slen, off, n_items));
    hipLaunchKernelGGL(file_bounds_kernel, dim3((F + 256) / 256), dim3(256), 0, ctx->stream, file, n_items, F,
                       (const uint64_t *)off, d_fpos, d_fbyte);
    MP_HIP(hipGetLastError());
    std::vector<uint64_t> fpos(F + 1), fbyte(F + 1);
    MP_HIP(hipMemcpyAsync(fpos.data(), d_fpos, sizeof(uint64_t) * (F + 1), hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipMemcpyAsync(fbyte.data(), d_fbyte, sizeof(uint64_t) * (F + 1), hipMemcpyDeviceToHost, ctx->stream));
    uint64_t total = 0;
    MP_HIP(hipMemcpyAsync(&total, off + n_items, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    MP_TRY(ensure(ctx, ctx->out, total + 64));
    const uint64_t real_items = fpos[F];
    if (real_items && total) {
        uint64_t *first = nullptr;
        MP_HIP(hipMallocAsync((void **)&first, ((total + kGBlock - 1) / kGBlock + 1) * sizeof(uint64_t), ctx->stream));
        const int rc = launch_gather(ctx, ctx->d_fq, ctx->fq_bytes, (const uint64_t *)off, real_items, total, total,
                                     first, RecSrc{(const uint64_t *)ctx->rec_off.p, slot},
                                     CopySink{(uint8_t *)ctx->out.p});
        MP_HIP(hipFreeAsync(first, ctx->stream));
        MP_TRY(rc);
    }
    hipEventRecord(ctx->ev[1], ctx->stream);
    MP_HIP(hipFreeAsync(d_small, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    if (bytes_out) *bytes_out = total;
    if (!out_prefix) return 0;
    // host side: one file per (t, P), written from pinned staging
    std::vector<char> host(total ? total : 1);
    if (total) MP_HIP(hipMemcpy(host.data(), ctx->out.p, total, hipMemcpyDeviceToHost));
    for (int f = 0; f < F; f++) {
        const int smp = f / (MP_NPART * T), p = (f / T) % MP_NPART, t = f % T;
        std::string dir = std::string(out_prefix) + (batch ? "/sample" + std::to_string(smp) : std::string()) + "/Rank0";
        if (batch && p == 0 && make_dirs(dir)) return ctx->fail(MP_EIO, "cannot create %s", dir.c_str());
        std::string name = dir + "/FromTid" + std::to_string(t) + "_ToRank" + std::to_string(p) + ".paired.fastq";
        FILE *fp = fopen(name.c_str(), "wb");
        if (!fp) return ctx->fail(MP_EIO, "cannot create %s", name.c_str());
        const uint64_t a = fbyte[f], b = fbyte[f + 1];
        if (b > a && fwrite(host.data() + a, 1, b - a, fp) != b - a) {
            fclose(fp);
            return ctx->fail(MP_EIO, "short write to %s", name.c_str());
        }
        fclose(fp);
    }
    return 0;
}

}  // namespace mp
