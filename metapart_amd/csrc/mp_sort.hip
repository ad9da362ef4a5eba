// Stable LSD radix sort of the scan's tuples by k-mer.
//
// The reference sorts 14-byte tuples with 9 stable byte passes per thread
// segment: byte 8 (adj) first, then k-mer bytes 0..7 (sortKmerList and
// radixPass, metapart.c:1494-1598, 2167-2198), after a range partition by
// segment.  Segments are contiguous bucket ranges of the k-mer's top bits, so
// one global stable sort gives the same array.  Stability keeps generation
// order among equal keys, as the reference's counting passes do.  Only the
// k-mer passes run: inside a run of equal k-mers the filter and the union do
// not depend on the order, and the few places that need the reference's adj
// order (the first-run quirk, the tuple dump) derive it from the run, which
// is in generation order.  The adj pass stays available (pass 0).
//
// MI355X design: structure-of-arrays tuples (keys u64, labels u32, adj u8;
// 13 bytes), one histogram read for every pass up front, then one "onesweep"
// kernel per 8-bit digit:
//   - 1024 lanes take a 6144-tuple tile (tile ids from an atomic counter, so
//     earlier tiles are always resident or done); two tiles fit a CU,
//   - each wave ranks its 384 tuples stably with 8 ballots per item
//     (bit-sliced digit match) and per-wave digit counters in LDS,
//   - per-digit tile counts (from the per-wave counters of the ranking) are
//     published and the digit's global prefix is found by decoupled look-back
//     over earlier tiles (epoch-tagged status words, so the status array is
//     never cleared between passes),
//   - tuples are reordered by digit in LDS and written out, so consecutive
//     lanes write consecutive addresses of each digit's run.
// Passes whose digit is constant over the scan (e.g. the top byte when K<29)
// are skipped: a stable pass on a constant digit is the identity.
//
// Batches of samples add a last pass whose digit is the sample of the tuple's
// label (labels are read ids, samples are contiguous id ranges): the result is
// every sample's tuples in (k-mer, adj) order, samples one after another, the
// same as sorting each sample alone.
#include <cstdlib>

#include "mp_internal.h"

namespace mp {

constexpr int kRadix = 256;
constexpr int kPasses = 10;  // adj, k-mer bytes 0..7, sample of the label

constexpr uint64_t kStAgg = 1, kStInc = 2;
constexpr uint64_t kStVal = (1ull << 46) - 1;

__device__ __forceinline__ uint64_t st_pack(uint32_t epoch, uint64_t flag, uint64_t v)
{
    return ((uint64_t)epoch << 48) | (flag << 46) | v;
}

struct SortPass {
    const uint64_t *keys_in;
    const uint32_t *lab_in;
    const uint8_t *adj_in;
    uint64_t *keys_out;
    uint32_t *lab_out;
    uint8_t *adj_out;
    uint64_t n;
    int shift;  // -1: the digit is the adj byte, -2: the sample of the label
    const uint64_t *grp_first;  // n_grp + 1 first read ids of the samples
    int n_grp;
    const uint64_t *digit_base;
    uint64_t *status;
    uint32_t *tile_ctr;
    uint32_t epoch;
};

// sample of a read id: last g with first[g] <= x
__device__ __forceinline__ uint32_t group_of(const uint64_t *first, int n, uint32_t x)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return (uint32_t)lo;
}

__global__ __launch_bounds__(kRadix) void sort_hist_kernel(const uint64_t *__restrict__ keys,
                                                           const uint8_t *__restrict__ adj,
                                                           const uint32_t *__restrict__ lab,
                                                           const uint64_t *__restrict__ grp_first, int n_grp,
                                                           uint64_t n, unsigned long long *__restrict__ ghist)
{
    __shared__ uint32_t h[kPasses][kRadix];
    for (int j = threadIdx.x; j < kPasses * kRadix; j += blockDim.x) (&h[0][0])[j] = 0;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t k = keys[i];
        atomicAdd(&h[0][adj[i]], 1u);
#pragma unroll
        for (int b = 0; b < 8; b++) atomicAdd(&h[1 + b][(k >> (8 * b)) & 0xFF], 1u);
        if (grp_first) atomicAdd(&h[9][group_of(grp_first, n_grp, lab[i])], 1u);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kPasses * kRadix; j += blockDim.x) {
        const uint32_t v = (&h[0][0])[j];
        if (v) atomicAdd(&ghist[j], (unsigned long long)v);
    }
}

__device__ __forceinline__ uint32_t digit_of(uint64_t k, uint8_t a, int shift)
{
    return shift < 0 ? (uint32_t)a : (uint32_t)((k >> shift) & 0xFF);
}

template <int kSB, int kSI>
__global__ __launch_bounds__(kSB) void onesweep_kernel(SortPass P)
{
    constexpr int kSW = kSB / kWave;
    constexpr int kTile = kSB * kSI;
    // one staging buffer (keys, then labels + adj) keeps LDS under half a CU,
    // so two tiles run per CU and one tile's memory phases overlap the other's
    // ranking and look-back
    __shared__ uint64_t s_buf[kTile];
    __shared__ uint8_t s_dig[kTile];
    __shared__ uint16_t s_cnt[kSW][kRadix];
    __shared__ uint32_t s_dstart[kRadix];  // digit starts in the tile
    __shared__ uint64_t s_gbase[kRadix];
    __shared__ uint32_t s_wsum[kRadix / kWave];
    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    const unsigned lane = __lane_id();
    for (int j = tid; j < kSW * kRadix; j += kSB) (&s_cnt[0][0])[j] = 0;
    if (tid == 0) s_tile = atomicAdd(P.tile_ctr, 1u);
    __syncthreads();
    const uint64_t tile = s_tile;
    const uint64_t base = tile * kTile;

    uint64_t k[kSI];
    uint32_t l[kSI];
    uint8_t a[kSI];
    uint32_t d[kSI];
    uint32_t rank[kSI];
#pragma unroll
    for (int i = 0; i < kSI; i++) {
        const uint64_t idx = base + (uint64_t)wave * (kSI * kWave) + (uint64_t)i * kWave + lane;
        if (idx < P.n) {
            k[i] = P.keys_in[idx];
            l[i] = P.lab_in[idx];
            a[i] = P.adj_in[idx];
            d[i] = P.shift == -2 ? group_of(P.grp_first, P.n_grp, l[i]) : digit_of(k[i], a[i], P.shift);
        } else {
            k[i] = 0;
            l[i] = 0;
            a[i] = 0;
            d[i] = kRadix;  // invalid
        }
    }
    // stable rank within the wave: item-major, lane-minor order
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int i = 0; i < kSI; i++) {
        const bool valid = d[i] < (uint32_t)kRadix;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d[i] >> b) & 1;
            const uint64_t m = __ballot(valid && bit);
            peers &= bit ? m : ~m;
        }
        rank[i] = 0;
        if (valid) {
            const uint32_t below = (uint32_t)__popcll(peers & lt);
            const uint32_t cnt = (uint32_t)__popcll(peers);
            const uint32_t old = s_cnt[wave][d[i]];
            rank[i] = old + below;
            __builtin_amdgcn_wave_barrier();
            if (below == cnt - 1) s_cnt[wave][d[i]] = (uint16_t)(old + cnt);
        }
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: exclusive offsets over waves, the tile's count (published for
    // later tiles' look-back), and the tile-local digit start
    uint32_t total = 0;
    if (tid < kRadix) {
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < kSW; w++) {
            const uint32_t c = s_cnt[w][tid];
            s_cnt[w][tid] = (uint16_t)run;
            run += c;
        }
        total = run;
        st_agent(P.status + tile * kRadix + tid, st_pack(P.epoch, tile == 0 ? kStInc : kStAgg, total));
        uint32_t inc = total;
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t v = __shfl_up(inc, o, kWave);
            if (lane >= (unsigned)o) inc += v;
        }
        if (lane == kWave - 1) s_wsum[wave] = inc;
        s_dstart[tid] = inc - total;  // wave-local exclusive prefix; waves added below
    }
    __syncthreads();
    if (tid < kRadix) {
        uint32_t woff = 0;
        for (int w = 0; w < wave; w++) woff += s_wsum[w];
        const uint32_t ds = s_dstart[tid] + woff;
        s_dstart[tid] = ds;
#pragma unroll
        for (int w = 0; w < kSW; w++) s_cnt[w][tid] = (uint16_t)(s_cnt[w][tid] + ds);  // tile position of wave w's first
        // decoupled look-back for digit tid
        uint64_t prefix = 0;
        if (tile != 0) {
            int64_t j = (int64_t)tile - 1;
            while (j >= 0) {
                const uint64_t s = ld_agent(P.status + (uint64_t)j * kRadix + tid);
                const uint64_t flag = (s >> 46) & 3;
                if ((uint32_t)(s >> 48) != P.epoch || flag == 0) continue;
                prefix += s & kStVal;
                if (flag == kStInc) break;
                j--;
            }
            st_agent(P.status + tile * kRadix + tid, st_pack(P.epoch, kStInc, prefix + total));
        }
        s_gbase[tid] = P.digit_base[tid] + prefix - ds;  // output index of tile position 0 of this digit
    }
    __syncthreads();
    // tile-local position of every item, then two staged write-outs
#pragma unroll
    for (int i = 0; i < kSI; i++) {
        if (d[i] < (uint32_t)kRadix) {
            rank[i] += s_cnt[wave][d[i]];
            s_buf[rank[i]] = k[i];
            s_dig[rank[i]] = (uint8_t)d[i];
        }
    }
    __syncthreads();
    const uint32_t nvalid = (uint32_t)(P.n - base < (uint64_t)kTile ? P.n - base : (uint64_t)kTile);
    for (uint32_t j = tid; j < nvalid; j += kSB) {
        P.keys_out[s_gbase[s_dig[j]] + j] = s_buf[j];
    }
    __syncthreads();
    uint32_t *s_lab = (uint32_t *)s_buf;
    uint8_t *s_adj = (uint8_t *)(s_lab + kTile);
#pragma unroll
    for (int i = 0; i < kSI; i++) {
        if (d[i] < (uint32_t)kRadix) {
            s_lab[rank[i]] = l[i];
            s_adj[rank[i]] = a[i];
        }
    }
    __syncthreads();
    for (uint32_t j = tid; j < nvalid; j += kSB) {
        const uint64_t g = s_gbase[s_dig[j]] + j;
        P.lab_out[g] = s_lab[j];
        P.adj_out[g] = s_adj[j];
    }
}

int radix_sort_generic(mp_ctx *ctx, DevArray *keys, DevArray *lab, DevArray *adj, int *cur, uint64_t n,
                       uint32_t pass_mask, int *passes_out, const uint64_t *d_grp_first, int n_grp,
                       uint64_t *grp_base)
{
    *passes_out = 0;
    if (n <= 1) return 0;
    // histograms of every digit
    MP_TRY(ensure(ctx, ctx->hist, sizeof(uint64_t) * (kPasses * kRadix + 64)));
    unsigned long long *gh = (unsigned long long *)ctx->hist.p;
    uint32_t *tile_ctr = (uint32_t *)(gh + kPasses * kRadix);
    MP_TRY(fill_u32(ctx, gh, 0, 2 * kPasses * kRadix + 2 * 32));
    hipEventRecord(ctx->pev[0], ctx->stream);
    {
        uint64_t blocks = (n + kRadix * 16 - 1) / (kRadix * 16);
        if (blocks > 2048) blocks = 2048;
        hipLaunchKernelGGL(sort_hist_kernel, dim3((unsigned)blocks), dim3(kRadix), 0, ctx->stream,
                           (const uint64_t *)keys[*cur].p, (const uint8_t *)adj[*cur].p,
                           (const uint32_t *)lab[*cur].p, d_grp_first, n_grp, n, gh);
        MP_HIP(hipGetLastError());
    }
    hipEventRecord(ctx->pev[1], ctx->stream);
    uint64_t h[kPasses * kRadix];
    MP_HIP(hipMemcpyAsync(h, gh, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    if (!d_grp_first) h[9 * kRadix] = n;  // no samples: one constant digit
    // digit base offsets of the passes that do work
    uint64_t base[kPasses][kRadix];
    bool run[kPasses];
    for (int p = 0; p < kPasses; p++) {
        uint64_t acc = 0, mx = 0;
        for (int r = 0; r < kRadix; r++) {
            base[p][r] = acc;
            acc += h[p * kRadix + r];
            if (h[p * kRadix + r] > mx) mx = h[p * kRadix + r];
        }
        if (acc != n)
            return ctx->fail(MP_EDEVICE, "sort histogram total %llu != %llu", (unsigned long long)acc,
                             (unsigned long long)n);
        run[p] = ((pass_mask >> p) & 1) && mx != n;
    }
    if (n_grp > 0 && grp_base) {  // tuple offset of every sample after the sort
        for (int g = 0; g < n_grp; g++) grp_base[g] = base[9][g];
        grp_base[n_grp] = n;
    }
    // tile shape: lanes x items.  Default from a sweep on MI355X at 2.4e9
    // tuples: 1024 x 6 (two tiles per CU, 22.8 ms/pass) < 1024 x 8 (one
    // tile per CU, 27.8) < 512 x 8 (32.8) < 512 x 16 (34.3).
    int shape = ctx->cfg.sort_tile;
    if (const char *e = getenv("MP_SORT_TILE")) shape = atoi(e);
    int SB = 1024, SI = 6;
    switch (shape) {
    case 4096: SB = 512; SI = 8; break;
    case 6144: SB = 512; SI = 12; break;
    case 8192: SB = 1024; SI = 8; break;
    default: SB = 1024; SI = 6; break;  // 6144 tuples on 1024 lanes
    }
    const uint64_t kTile = (uint64_t)SB * SI;
    const uint64_t tiles = (n + kTile - 1) / kTile;
    const uint64_t st_bytes = tiles * kRadix * sizeof(uint64_t);
    if (ctx->status.bytes < st_bytes) {
        MP_TRY(ensure(ctx, ctx->status, st_bytes));
        MP_TRY(fill_u32(ctx, ctx->status.p, 0, ctx->status.bytes / 4));
        ctx->epoch = 0;
    }
    uint64_t *d_base = nullptr;
    MP_HIP(hipMallocAsync((void **)&d_base, sizeof(base), ctx->stream));
    MP_HIP(hipMemcpyAsync(d_base, base, sizeof(base), hipMemcpyHostToDevice, ctx->stream));
    int executed = 0;
    for (int p = 0; p < kPasses; p++) {
        if (!run[p]) continue;
        if (++ctx->epoch == 0) {  // wrapped: old status words could alias
            MP_TRY(fill_u32(ctx, ctx->status.p, 0, ctx->status.bytes / 4));
            ctx->epoch = 1;
        }
        const int src = *cur, dst = 1 - *cur;
        SortPass P;
        P.keys_in = (const uint64_t *)keys[src].p;
        P.lab_in = (const uint32_t *)lab[src].p;
        P.adj_in = (const uint8_t *)adj[src].p;
        P.keys_out = (uint64_t *)keys[dst].p;
        P.lab_out = (uint32_t *)lab[dst].p;
        P.adj_out = (uint8_t *)adj[dst].p;
        P.n = n;
        P.shift = p == 0 ? -1 : p == 9 ? -2 : 8 * (p - 1);
        P.grp_first = d_grp_first;
        P.n_grp = n_grp;
        P.digit_base = d_base + p * kRadix;
        P.status = (uint64_t *)ctx->status.p;
        P.tile_ctr = tile_ctr + executed;
        P.epoch = ctx->epoch;
        hipEventRecord(ctx->pev[2 + 2 * executed], ctx->stream);
        if (SB == 512 && SI == 8)
            hipLaunchKernelGGL((onesweep_kernel<512, 8>), dim3((unsigned)tiles), dim3(512), 0, ctx->stream, P);
        else if (SB == 512 && SI == 12)
            hipLaunchKernelGGL((onesweep_kernel<512, 12>), dim3((unsigned)tiles), dim3(512), 0, ctx->stream, P);
        else if (SB == 1024 && SI == 8)
            hipLaunchKernelGGL((onesweep_kernel<1024, 8>), dim3((unsigned)tiles), dim3(1024), 0, ctx->stream, P);
        else
            hipLaunchKernelGGL((onesweep_kernel<1024, 6>), dim3((unsigned)tiles), dim3(1024), 0, ctx->stream, P);
        MP_HIP(hipGetLastError());
        hipEventRecord(ctx->pev[3 + 2 * executed], ctx->stream);
        *cur = dst;
        executed++;
    }
    MP_HIP(hipFreeAsync(d_base, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    float tot = 0, ms = 0;
    for (int e = 0; e < executed; e++) {
        hipEventElapsedTime(&ms, ctx->pev[2 + 2 * e], ctx->pev[3 + 2 * e]);
        tot += ms;
    }
    ctx->sort_pass_avg = executed ? tot / executed : 0.f;
    hipEventElapsedTime(&ctx->sort_hist_ms, ctx->pev[0], ctx->pev[1]);
    *passes_out = executed;
    return 0;
}

int radix_sort_tuples(mp_ctx *ctx, uint64_t n, int *passes_out)
{
    // k-mer bytes only: within a run nothing downstream depends on the adj
    // order except the quirk's first tuples and the tuple dump, which derive it
    return radix_sort_generic(ctx, ctx->keys, ctx->lab, ctx->adj, &ctx->cur, n, 0x1FE, passes_out, nullptr, 0,
                              nullptr);
}

}  // namespace mp
