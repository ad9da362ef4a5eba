// Device memory accounting, fills and a single-pass (decoupled look-back)
// exclusive scan used for record, k-mer and output offsets.
#include "mp_internal.h"

namespace mp {

// a region of the ctx's arena (arena mode: nothing is allocated or freed;
// the arena is reset as a whole at the start of each call that uses it)
static int arena_carve(mp_ctx *ctx, void **p, uint64_t bytes)
{
    const uint64_t off = (ctx->arena_used + 255) & ~255ull;
    if (off + bytes > ctx->arena_bytes)
        return ctx->fail(MP_ENOMEM, "the input needs more than the %llu-byte fixed scratch (%llu bytes in use, %llu "
                                    "more asked)", (unsigned long long)ctx->arena_bytes,
                         (unsigned long long)ctx->arena_used, (unsigned long long)bytes);
    *p = ctx->arena + off;
    ctx->arena_used = off + bytes;
    return 0;
}

int tmp_alloc(mp_ctx *ctx, void **p, uint64_t bytes)
{
    if (ctx->arena) return arena_carve(ctx, p, bytes);
    MP_HIP(hipMallocAsync(p, bytes, ctx->stream));
    return 0;
}

int tmp_free(mp_ctx *ctx, void *p)
{
    if (ctx->arena || !p) return 0;
    MP_HIP(hipFreeAsync(p, ctx->stream));
    return 0;
}

int ensure(mp_ctx *ctx, DevArray &a, uint64_t bytes)
{
    if (bytes == 0) bytes = 256;
    if (a.bytes >= bytes) return 0;
    if (ctx->arena) {  // the old region stays carved until the arena is reset
        MP_TRY(arena_carve(ctx, &a.p, bytes));
        a.bytes = bytes;
        return 0;
    }
    uint64_t want = bytes + (bytes >> 4);  // 6 % slack so slowly growing runs do not reallocate each time
    // a buffer that grows again at least doubles while a quarter of the budget
    // stays free: a stream of ever larger inputs (a served job mix) then pays a
    // few reallocations (a device synchronisation, a free and an allocation of
    // gigabytes each), not one per new largest input
    if (a.bytes && 2 * a.bytes > want && ctx->held - a.bytes + 2 * a.bytes <= ctx->budget / 4 * 3) want = 2 * a.bytes;
    if (ctx->held - a.bytes + want > ctx->budget) {
        want = bytes;
        if (ctx->held - a.bytes + want > ctx->budget)
            return ctx->fail(MP_ENOMEM, "device budget exceeded: need %llu more bytes (held %llu of %llu)",
                             (unsigned long long)want, (unsigned long long)ctx->held, (unsigned long long)ctx->budget);
    }
    if (a.p) {
        MP_HIP(hipStreamSynchronize(ctx->stream));
        MP_HIP(hipFree(a.p));
        ctx->held -= a.bytes;
        a.p = nullptr;
        a.bytes = 0;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return ctx->fail(MP_ENOMEM, "hipMalloc(%llu) failed: %s", (unsigned long long)want, hipGetErrorString(e));
    }
    a.p = p;
    a.bytes = want;
    ctx->held += want;
    return 0;
}

void release(mp_ctx *ctx, DevArray &a)
{
    if (a.p && !ctx->arena) {
        (void)hipFree(a.p);
        ctx->held -= a.bytes;
    }
    a.p = nullptr;
    a.bytes = 0;
}

__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, uint64_t n)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

int fill_u32(mp_ctx *ctx, void *p, uint32_t v, uint64_t n_words)
{
    if (n_words == 0) return 0;
    uint64_t blocks = (n_words + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, (uint32_t *)p, v, n_words);
    MP_HIP(hipGetLastError());
    return 0;
}

__global__ void iota_u32_kernel(uint32_t *p, uint64_t n)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}

int iota_u32(mp_ctx *ctx, uint32_t *p, uint64_t n)
{
    if (n == 0) return 0;
    hipLaunchKernelGGL(iota_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, p, n);
    MP_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- scan
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;
constexpr uint64_t kFlagAgg = 1ull << 62, kFlagInc = 2ull << 62, kValMask = (1ull << 62) - 1;

template <typename T>
__global__ __launch_bounds__(kScanBlock) void scan_kernel(const T *__restrict__ in, uint64_t *__restrict__ out,
                                                          uint64_t n, uint64_t *status, uint32_t *tile_ctr)
{
    __shared__ uint64_t s_wave[kScanBlock / kWave];
    __shared__ uint64_t s_prefix;
    __shared__ uint32_t s_tile;
    if (threadIdx.x == 0) s_tile = atomicAdd(tile_ctr, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t base = (uint64_t)tile * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint64_t v[kScanItems];
    uint64_t sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        uint64_t idx = base + i;
        v[i] = idx < n ? (uint64_t)in[idx] : 0;
        sum += v[i];
    }
    // inclusive scan of per-thread sums inside the wave
    const unsigned lane = __lane_id();
    uint64_t inc = sum;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint64_t o = __shfl_up(inc, d, kWave);
        if (lane >= (unsigned)d) inc += o;
    }
    const int wave = threadIdx.x / kWave;
    if (lane == kWave - 1) s_wave[wave] = inc;
    __syncthreads();
    uint64_t wave_off = 0, block_total = 0;
#pragma unroll
    for (int w = 0; w < kScanBlock / kWave; w++) {
        if (w < wave) wave_off += s_wave[w];
        block_total += s_wave[w];
    }
    if (threadIdx.x == 0) {
        uint64_t prefix = 0;
        if (tile == 0) {
            st_agent(&status[0], kFlagInc | block_total);
        } else {
            st_agent(&status[tile], kFlagAgg | block_total);
            int64_t j = (int64_t)tile - 1;
            while (j >= 0) {
                uint64_t s = ld_agent(&status[j]);
                uint64_t flag = s & ~kValMask;
                if (flag == 0) continue;
                prefix += s & kValMask;
                if (flag == kFlagInc) break;
                j--;
            }
            st_agent(&status[tile], kFlagInc | (prefix + block_total));
        }
        s_prefix = prefix;
    }
    __syncthreads();
    uint64_t run = s_prefix + wave_off + inc - sum;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        uint64_t idx = base + i;
        if (idx < n) out[idx] = run;
        run += v[i];
    }
    if (base < n && base + kScanItems >= n) out[n] = run;  // thread holding the last element
}

template <typename T>
static int scan_impl(mp_ctx *ctx, const T *in, uint64_t *out, uint64_t n)
{
    if (n == 0) {
        return fill_u32(ctx, out, 0, 2);
    }
    uint64_t tiles = (n + kScanTile - 1) / kScanTile;
    MP_TRY(ensure(ctx, ctx->scanst, (tiles + 64) * sizeof(uint64_t)));
    uint64_t *status = (uint64_t *)ctx->scanst.p;
    uint32_t *ctr = (uint32_t *)(status + tiles);
    MP_TRY(fill_u32(ctx, status, 0, tiles * 2 + 2));
    hipLaunchKernelGGL(scan_kernel<T>, dim3((unsigned)tiles), dim3(kScanBlock), 0, ctx->stream, in, out, n, status, ctr);
    MP_HIP(hipGetLastError());
    return 0;
}

int scan_u32(mp_ctx *ctx, const uint32_t *in, uint64_t *out, uint64_t n) { return scan_impl(ctx, in, out, n); }
int scan_u64(mp_ctx *ctx, const uint64_t *in, uint64_t *out, uint64_t n) { return scan_impl(ctx, in, out, n); }

}  // namespace mp
