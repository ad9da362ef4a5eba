// The allocation-free m-mer histogram entry point (mp_count_chunks_noalloc).
//
// Same result as mp_count_chunks for FASTQ bytes the caller already holds in
// device memory, for callers that may not allocate device memory (tight
// request loops, harnesses that forbid it).  Every device buffer of the count
// path -- newline counts, the record table, the 2-bit and 'N' streams, the
// histogram's tile tables and low-bit staging, the per-chunk counts -- is
// carved from one fixed array in device memory, `scratch`, which each GPU
// the library runs on holds until the process ends.  Each call starts the
// carving again from the start of the array, so calls that overlap on one
// GPU race; a call clears every counter it reads before using it.
//
// Set-up, once per device on its first call: the device address of the
// array (hipGetSymbolAddress: in host code the array's name is not a device
// address) and a context (stream and events) whose buffers come from it.
#include <cstring>
#include <mutex>

#include "mp_internal.h"

using namespace mp;

namespace {

constexpr uint64_t kScratchBytes = MP_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

struct DeviceState {
    uint8_t *ptr = nullptr;  // device address of scratch on this device
    mp_ctx *ctx = nullptr;   // arena-mode context over it
};
DeviceState g_dev[kMaxDevices];
std::mutex g_setup;

int setup(int device, DeviceState **out)
{
    std::lock_guard<std::mutex> lk(g_setup);
    DeviceState &d = g_dev[device];
    if (!d.ptr) {
        if (hipSetDevice(device) != hipSuccess) return MP_EDEVICE;
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return MP_EDEVICE;
        }
        d.ptr = (uint8_t *)p;
    }
    if (!d.ctx) {
        mp_cfg cfg;
        memset(&cfg, 0, sizeof(cfg));
        cfg.device = device;
        cfg.k = 31;
        cfg.hbm_budget = kScratchBytes;
        mp_ctx *ctx = nullptr;
        const int rc = mp_ctx_create(&cfg, &ctx);
        if (rc) return rc;
        ctx->arena = d.ptr;
        ctx->arena_bytes = kScratchBytes;
        d.ctx = ctx;
    }
    *out = &d;
    return 0;
}

}  // namespace

extern "C" {

int mp_count_chunks_noalloc(int device, int k, int min3, const void *d_fq, uint64_t nbytes,
                            const uint64_t *chunk_ofs, int C, uint32_t *counts, uint64_t *reads_per_chunk,
                            uint64_t *total_kmers)
{
    if (!d_fq || !chunk_ofs || !counts || C <= 0 || k <= 10 || k > 32) return MP_EINVAL;
    if (nbytes > MP_NOALLOC_MAX_FASTQ || C > MP_NOALLOC_MAX_CHUNKS) return MP_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev || device >= kMaxDevices) {
        (void)hipGetLastError();
        return MP_EINVAL;
    }
    DeviceState *d = nullptr;
    MP_TRY(setup(device, &d));
    mp_ctx *ctx = d->ctx;
    if (hipSetDevice(device) != hipSuccess) return MP_EDEVICE;
    // every buffer is carved again from the start of the array
    ctx->for_each_array([](DevArray &a) {
        a.p = nullptr;
        a.bytes = 0;
    });
    ctx->arena_used = 0;
    ctx->err.clear();
    ctx->cfg.k = k;
    ctx->cfg.min3 = min3 ? 1 : 0;
    ctx->d_fq = (const uint8_t *)d_fq;
    ctx->fq_owned = false;
    ctx->fq_bytes = nbytes;
    ctx->N = 0;
    ctx->batch = false;
    ctx->have_scan = ctx->finished = ctx->khist_valid = false;
    ctx->use_labsnap = false;
    ctx->scans_done = 0;
    if (chunk_ofs[0] != 0 || chunk_ofs[C] != nbytes) return ctx->fail(MP_EINVAL, "chunks must cover the FASTQ bytes");
    for (int c = 0; c < C; c++)
        if (chunk_ofs[c + 1] < chunk_ofs[c]) return ctx->fail(MP_EINVAL, "chunk offsets must not decrease");
    MP_TRY(parse_reads(ctx, chunk_ofs, nullptr, C));
    const uint64_t words = (uint64_t)C * (MP_NBUCKETS + 2);
    MP_TRY(ensure(ctx, ctx->scratch, words * sizeof(uint32_t)));
    uint32_t *d_counts = (uint32_t *)ctx->scratch.p;
    MP_TRY(count_chunks_device(ctx, ctx->chunk_ids.data(), C, d_counts));
    MP_HIP(hipMemcpyAsync(counts, d_counts, words * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    uint64_t tot = 0;
    for (uint64_t i = 0; i < words; i++) tot += counts[i];
    if (total_kmers) *total_kmers = tot;
    if (reads_per_chunk)
        for (int c = 0; c < C; c++) reads_per_chunk[c] = ctx->chunk_ids[c + 1] - ctx->chunk_ids[c];
    return 0;
}

}  // extern "C"
