// Canonical k-mer kernels: m-mer histograms and ordered tuple enumeration.
//
// Plain variant (every k-mer; countMmers.c:83-217, enumerateKmers.c:127-371).
// The reference rolls a forward and a reverse-complement code through each
// read on one CPU thread.  Here a workgroup owns 64 consecutive reads and its
// 256 lanes walk the flattened list of their k-mer positions, 256 at a time,
// so every lane does O(1) work per k-mer regardless of read length:
//   F  = 2K bits of the packed stream at the k-mer's first base (two 64-bit
//        loads and a funnel shift; MSB-first packing makes this the
//        reference's curKmerInt),
//   R  = reverse complement of F by bit reversal (the reference's curRcInt),
//   N  = the same window of the 1-bit 'N' stream must be empty (lastNPos),
//   canonical = F < R ? F : R, bucket = canonical >> 2(K-10) (getMmerPrefix).
// At K = 32 the reference's mask (1 << 64) - 1 evaluates to 0 on x86, so every
// forward code after the first position is 0; that is reproduced.
// Emission is an ordered stream compaction (wave ballots + per-workgroup
// offsets from a count pass), so the tuple array is in the reference's
// generation order: read order, then position order within a read.
//
// 3-minimizer variant (countMmers_3minimizers.c:83-271,
// enumerateKmers_3minimizers.c:129-474): a sequential window state machine per
// read, one lane per read, restating the reference loop on the raw bytes.
#include "mp_internal.h"

namespace mp {

constexpr int kRG = 64;       // reads per workgroup, plain kernels
constexpr int kKBlock = 256;  // lanes per workgroup

struct KParams {
    const uint8_t *fq;
    const uint64_t *seq_off;
    const uint32_t *seq_len;
    const uint64_t *base_off;
    const uint64_t *pos_off;
    const uint64_t *packed;
    const uint64_t *nmask;
    uint64_t N;
    int K;
    int bshift;
    uint64_t lo, hi;
    // histogram
    uint32_t *counts;
    const uint64_t *chunk_ids;
    int C;
    // count / emit
    uint32_t *wg_cnt;
    const uint64_t *wg_off;
    uint64_t *keys;
    uint8_t *adj;
    uint32_t *lab;
    const uint32_t *parent;
    const uint8_t *bad;
    int scan0;
};

enum { kModeHist = 0, kModeCount = 1, kModeEmit = 2 };

__device__ __forceinline__ uint8_t rc_char(uint8_t c)
{
    // "TGAC   N"[(c >> 1) & 7]  (enumerateKmers.c:237)
    return (uint8_t)((0x4E20202043414754ull >> (8 * ((c >> 1) & 7))) & 0xFF);
}
__device__ __forceinline__ uint8_t adj_of(uint8_t prev, uint8_t next)
{
    return (uint8_t)(((prev << 2) & 0xF8) | ((next >> 1) & 0x7));
}
__device__ __forceinline__ uint32_t code_of(uint8_t c) { return (0x2310u >> (((c & 6) >> 1) * 4)) & 3; }

// Canonical k-mer of the window starting at base b (position p of a read with
// sequence at fq[so], length L).  Returns false if the window holds an 'N'.
__device__ __forceinline__ bool kmer_at(const KParams &P, uint64_t b, uint32_t p, uint64_t &canon, bool &fwd)
{
    const int K = P.K;
    {
        const uint64_t w = b >> 6;
        const int o = (int)(b & 63);
        const uint64_t hi = P.nmask[w], lo = P.nmask[w + 1];
        const int s = o + K;
        const uint64_t v = s <= 64 ? (hi >> (64 - s)) : ((hi << (s - 64)) | (lo >> (128 - s)));
        if (v & ((1ull << K) - 1)) return false;
    }
    const uint64_t w = b >> 5;
    const int o = (int)(b & 31);
    const uint64_t hi = P.packed[w], lo = P.packed[w + 1];
    const int s = 2 * (o + K);
    uint64_t F = s <= 64 ? (hi >> (64 - s)) : ((hi << (s - 64)) | (lo >> (128 - s)));
    if (K < 32) F &= (1ull << (2 * K)) - 1;
    const uint64_t R = revcomp(F, K);
    if (K == 32 && p > 0) F = 0;
    fwd = F < R;
    canon = fwd ? F : R;
    return true;
}

template <int MODE>
__global__ __launch_bounds__(kKBlock) void kmer_plain_kernel(KParams P)
{
    __shared__ uint32_t s_pref[kRG + 1];
    __shared__ uint64_t s_base[kRG], s_so[kRG];
    __shared__ uint32_t s_len[kRG], s_aux[kRG];
    __shared__ uint32_t s_wc[kKBlock / kWave];
    const uint64_t r0 = (uint64_t)blockIdx.x * kRG;
    const int nr = (int)((P.N - r0) < (uint64_t)kRG ? (P.N - r0) : (uint64_t)kRG);
    const int tid = threadIdx.x;
    const uint64_t p0 = P.pos_off[r0];
    if (tid <= nr) s_pref[tid] = (uint32_t)(P.pos_off[r0 + tid] - p0);
    if (tid < nr) {
        const uint64_t r = r0 + tid;
        s_base[tid] = P.base_off[r];
        s_so[tid] = P.seq_off[r];
        s_len[tid] = P.seq_len[r];
        if (MODE == kModeHist) {
            int lo = 0, hi = P.C - 1;  // chunk: last c with chunk_ids[c] <= r
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (P.chunk_ids[mid] <= r) lo = mid; else hi = mid - 1;
            }
            s_aux[tid] = (uint32_t)lo;
        } else if (MODE == kModeEmit) {
            s_aux[tid] = P.scan0 ? (uint32_t)r : (P.bad[r] ? (uint32_t)r : P.parent[r]);
        }
    }
    __syncthreads();
    const uint32_t total = s_pref[nr];
    uint32_t count = 0;
    uint64_t out = MODE == kModeEmit ? P.wg_off[blockIdx.x] : 0;
    const int wave = tid / kWave;
    for (uint32_t q0 = 0; q0 < total; q0 += kKBlock) {
        const uint32_t q = q0 + tid;
        bool v = false, fwd = false;
        uint64_t canon = 0;
        int rl = 0;
        uint32_t p = 0;
        if (q < total) {
            int lo = 0, hi = nr - 1;
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (s_pref[mid] <= q) lo = mid; else hi = mid - 1;
            }
            rl = lo;
            p = q - s_pref[rl];
            v = kmer_at(P, s_base[rl] + p, p, canon, fwd);
            if (MODE != kModeHist && v) {
                const uint64_t bk = canon >> P.bshift;
                v = bk >= P.lo && bk < P.hi;
            }
        }
        if (MODE == kModeHist) {
            if (v) atomicAdd(&P.counts[(size_t)s_aux[rl] * (MP_NBUCKETS + 2) + (canon >> P.bshift)], 1u);
        } else if (MODE == kModeCount) {
            count += v;
        } else {
            const uint64_t m = __ballot(v);
            if (__lane_id() == 0) s_wc[wave] = (uint32_t)__popcll(m);
            __syncthreads();
            uint32_t woff = 0, tot = 0;
#pragma unroll
            for (int w = 0; w < kKBlock / kWave; w++) {
                woff += w < wave ? s_wc[w] : 0;
                tot += s_wc[w];
            }
            if (v) {
                const uint64_t pos = out + woff + (uint64_t)__popcll(m & lanemask_lt());
                const uint64_t so = s_so[rl];
                const uint32_t L = s_len[rl];
                const int K = P.K;
                const uint8_t prev = p == 0 ? (uint8_t)'N' : P.fq[so + p - 1];
                const uint8_t next = p == L - K ? (uint8_t)'N' : P.fq[so + p + K];
                const uint8_t a = fwd ? adj_of(prev, next) : adj_of(rc_char(next), rc_char(prev));
                P.keys[pos] = canon;
                P.adj[pos] = a;
                P.lab[pos] = s_aux[rl];
            }
            out += tot;
            __syncthreads();
        }
    }
    if (MODE == kModeCount) {
        for (int d = kWave / 2; d > 0; d >>= 1) count += __shfl_down(count, d, kWave);
        if (__lane_id() == 0) s_wc[wave] = count;
        __syncthreads();
        if (tid == 0) {
            uint32_t t = 0;
            for (int w = 0; w < kKBlock / kWave; w++) t += s_wc[w];
            P.wg_cnt[blockIdx.x] = t;
        }
    }
}

// ---------------------------------------------------------------- 3-minimizer
// Rolling codes exactly as the reference (countMmers_3minimizers.c:107-190).
struct Roll {
    uint64_t F = 0, R = 0, mask = 0;
    int lastN = -1;
};

// counter: countMmers_3minimizers.c:83-271 (no range filter)
template <typename Sink>
__device__ void min3_count_read(const uint8_t *s, int L, int K, int bshift, Sink sink)
{
    if (L <= K) return;
    const int last = L - K;
    const uint64_t mask = 2 * K >= 64 ? 0ull : (1ull << (2 * K)) - 1;
    uint64_t F = 0, R = 0, minK = 0, minB = 0;
    int lastN = -1, next = last / 3;
    for (int j = 0; j < K; j++) {
        const uint8_t c = s[j];
        if (c == 'N') lastN = j;
        R |= (uint64_t)(3 - code_of(c)) << (2 * j);
        F = (F << 2) | code_of(c);
    }
    if (lastN == -1) {
        const uint64_t c = F < R ? F : R;
        minK = c;
        minB = c >> bshift;
    }
    for (int i = 1; i <= last; i++) {
        const uint8_t ch = s[i + K - 1];
        R = (R >> 2) | ((uint64_t)(3 - code_of(ch)) << (2 * (K - 1)));
        F = ((F << 2) | code_of(ch)) & mask;
        if (ch == 'N') lastN = i + K - 1;
        if (i > lastN) {
            const uint64_t c = F < R ? F : R;
            const uint64_t b = c >> bshift;
            if (i > next) {
                if (minK != 0) sink(minB);
                minK = c;
                minB = b;
                next += last / 3;
                if (next + last / 3 > last) next = last;
            } else if (c < minK || minK == 0) {
                minK = c;
                minB = b;
            }
        }
    }
    if (minK != 0) sink(minB);
}

// enumerator: enumerateKmers_3minimizers.c:129-474 (range filter before the window logic)
template <typename Sink>
__device__ void min3_enum_read(const uint8_t *s, int L, int K, int bshift, uint64_t lo, uint64_t hi, Sink sink)
{
    if (L <= K) return;
    const int last = L - K;
    const uint64_t mask = 2 * K >= 64 ? 0ull : (1ull << (2 * K)) - 1;
    uint64_t F = 0, R = 0, minK = 0;
    uint8_t minPrev = 'N', minNext = 'N';
    int lastN = -1;
    uint8_t prev = 'N', next = s[K];
    for (int j = 0; j < K; j++) {
        const uint8_t c = s[j];
        if (c == 'N') lastN = j;
        R |= (uint64_t)(3 - code_of(c)) << (2 * j);
        F = (F << 2) | code_of(c);
    }
    if (lastN == -1) {
        const bool fwd = F < R;
        const uint64_t c = fwd ? F : R;
        const uint64_t b = c >> bshift;
        if (b >= lo && b < hi) {
            minK = c;
            minPrev = fwd ? prev : rc_char(prev);
            minNext = fwd ? next : rc_char(next);
        }
    }
    int nextW = last / 3;
    for (int i = 1; i <= last; i++) {
        prev = s[i - 1];
        next = i == last ? (uint8_t)'N' : s[i + K];
        const uint8_t ch = s[i + K - 1];
        R = (R >> 2) | ((uint64_t)(3 - code_of(ch)) << (2 * (K - 1)));
        F = ((F << 2) | code_of(ch)) & mask;
        if (ch == 'N') lastN = i + K - 1;
        if (i <= lastN) continue;
        const bool fwd = F < R;
        const uint64_t c = fwd ? F : R;
        const uint64_t b = c >> bshift;
        if (!(b >= lo && b < hi)) continue;
        const uint8_t cp = fwd ? prev : rc_char(prev), cn = fwd ? next : rc_char(next);
        if (i > nextW) {
            if (minK != 0) sink(minK, fwd ? adj_of(minPrev, minNext) : adj_of(minNext, minPrev));
            minK = c;
            minPrev = cp;
            minNext = cn;
            nextW += last / 3;
            if (nextW + last / 3 > last) nextW = last;
        } else if (c < minK || minK == 0) {
            minK = c;
            minPrev = cp;
            minNext = cn;
        }
    }
    if (minK != 0) sink(minK, adj_of(minNext, minPrev));
}

template <int MODE>
__global__ __launch_bounds__(kKBlock) void kmer_min3_kernel(KParams P)
{
    __shared__ uint32_t s_wc[kKBlock / kWave];
    const uint64_t r = (uint64_t)blockIdx.x * kKBlock + threadIdx.x;
    const int wave = threadIdx.x / kWave;
    const unsigned lane = __lane_id();
    uint32_t cnt = 0;
    const uint8_t *s = nullptr;
    int L = 0;
    if (r < P.N) {
        s = P.fq + P.seq_off[r];
        L = (int)P.seq_len[r];
    }
    if (MODE == kModeHist) {
        if (r < P.N) {
            int lo = 0, hi = P.C - 1;
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (P.chunk_ids[mid] <= r) lo = mid; else hi = mid - 1;
            }
            uint32_t *cnts = P.counts + (size_t)lo * (MP_NBUCKETS + 2);
            min3_count_read(s, L, P.K, P.bshift, [&](uint64_t b) { atomicAdd(&cnts[b], 1u); });
        }
        return;
    }
    if (r < P.N) min3_enum_read(s, L, P.K, P.bshift, P.lo, P.hi, [&](uint64_t, uint8_t) { cnt++; });
    // exclusive scan of cnt over the workgroup
    uint32_t inc = cnt;
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t o = __shfl_up(inc, d, kWave);
        if (lane >= (unsigned)d) inc += o;
    }
    if (lane == kWave - 1) s_wc[wave] = inc;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kKBlock / kWave; w++) {
        woff += w < wave ? s_wc[w] : 0;
        tot += s_wc[w];
    }
    if (MODE == kModeCount) {
        if (threadIdx.x == 0) P.wg_cnt[blockIdx.x] = tot;
        return;
    }
    if (r >= P.N || cnt == 0) return;
    uint64_t pos = P.wg_off[blockIdx.x] + woff + inc - cnt;
    const uint32_t label = P.scan0 ? (uint32_t)r : (P.bad[r] ? (uint32_t)r : P.parent[r]);
    min3_enum_read(s, L, P.K, P.bshift, P.lo, P.hi, [&](uint64_t k, uint8_t a) {
        P.keys[pos] = k;
        P.adj[pos] = a;
        P.lab[pos] = label;
        pos++;
    });
}

__global__ void bucket_total_kernel(const uint32_t *__restrict__ counts, int C, uint64_t *__restrict__ tot)
{
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= MP_NBUCKETS) return;
    uint64_t s = 0;
    for (int c = 0; c < C; c++) s += counts[(size_t)c * (MP_NBUCKETS + 2) + b];
    tot[b] = s;
}

int bucket_totals(mp_ctx *ctx, const uint32_t *d_counts, int C, uint64_t *d_tot)
{
    hipLaunchKernelGGL(bucket_total_kernel, dim3(MP_NBUCKETS / 256), dim3(256), 0, ctx->stream, d_counts, C, d_tot);
    MP_HIP(hipGetLastError());
    return 0;
}

static KParams base_params(mp_ctx *ctx)
{
    KParams P{};
    P.fq = ctx->d_fq;
    P.seq_off = (const uint64_t *)ctx->seq_off.p;
    P.seq_len = (const uint32_t *)ctx->seq_len.p;
    P.base_off = (const uint64_t *)ctx->base_off.p;
    P.pos_off = (const uint64_t *)ctx->pos_off.p;
    P.packed = (const uint64_t *)ctx->packed.p;
    P.nmask = (const uint64_t *)ctx->nmask.p;
    P.N = ctx->N;
    P.K = ctx->cfg.k;
    P.bshift = 2 * (ctx->cfg.k - 10);
    return P;
}

int count_chunks_device(mp_ctx *ctx, const uint64_t *chunk_first_id, int C, uint32_t *d_counts)
{
    KParams P = base_params(ctx);
    uint64_t *d_ids = nullptr;
    MP_HIP(hipMallocAsync((void **)&d_ids, sizeof(uint64_t) * (C + 1), ctx->stream));
    MP_HIP(hipMemcpyAsync(d_ids, chunk_first_id, sizeof(uint64_t) * (C + 1), hipMemcpyHostToDevice, ctx->stream));
    MP_TRY(fill_u32(ctx, d_counts, 0, (uint64_t)C * (MP_NBUCKETS + 2)));
    P.counts = d_counts;
    P.chunk_ids = d_ids;
    P.C = C;
    if (ctx->cfg.min3) {
        hipLaunchKernelGGL(kmer_min3_kernel<kModeHist>, dim3((unsigned)((ctx->N + kKBlock - 1) / kKBlock)),
                           dim3(kKBlock), 0, ctx->stream, P);
    } else {
        hipLaunchKernelGGL(kmer_plain_kernel<kModeHist>, dim3((unsigned)((ctx->N + kRG - 1) / kRG)), dim3(kKBlock), 0,
                           ctx->stream, P);
    }
    MP_HIP(hipGetLastError());
    MP_HIP(hipFreeAsync(d_ids, ctx->stream));
    return 0;
}

int enumerate_scan(mp_ctx *ctx, int s, uint64_t lo, uint64_t hi, uint64_t *n_out)
{
    KParams P = base_params(ctx);
    P.lo = lo;
    P.hi = hi;
    P.scan0 = s == 0;
    P.parent = (const uint32_t *)ctx->parent.p;
    P.bad = (const uint8_t *)ctx->bad.p;
    const bool m3 = ctx->cfg.min3 != 0;
    const uint64_t per = m3 ? kKBlock : kRG;
    const uint64_t nwg = (ctx->N + per - 1) / per;
    MP_TRY(ensure(ctx, ctx->wg_cnt, nwg * sizeof(uint32_t)));
    MP_TRY(ensure(ctx, ctx->wg_off, (nwg + 1) * sizeof(uint64_t)));
    P.wg_cnt = (uint32_t *)ctx->wg_cnt.p;
    P.wg_off = (const uint64_t *)ctx->wg_off.p;
    if (m3) hipLaunchKernelGGL(kmer_min3_kernel<kModeCount>, dim3((unsigned)nwg), dim3(kKBlock), 0, ctx->stream, P);
    else hipLaunchKernelGGL(kmer_plain_kernel<kModeCount>, dim3((unsigned)nwg), dim3(kKBlock), 0, ctx->stream, P);
    MP_HIP(hipGetLastError());
    MP_TRY(scan_u32(ctx, (const uint32_t *)ctx->wg_cnt.p, (uint64_t *)ctx->wg_off.p, nwg));
    uint64_t n = 0;
    MP_HIP(hipMemcpyAsync(&n, (uint64_t *)ctx->wg_off.p + nwg, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    for (int b = 0; b < 2; b++) {
        MP_TRY(ensure(ctx, ctx->keys[b], (n + 1) * sizeof(uint64_t)));
        MP_TRY(ensure(ctx, ctx->lab[b], (n + 1) * sizeof(uint32_t)));
        MP_TRY(ensure(ctx, ctx->adj[b], (n + 16) * sizeof(uint8_t)));
    }
    P.keys = (uint64_t *)ctx->keys[0].p;
    P.adj = (uint8_t *)ctx->adj[0].p;
    P.lab = (uint32_t *)ctx->lab[0].p;
    if (n > 0) {
        if (m3) hipLaunchKernelGGL(kmer_min3_kernel<kModeEmit>, dim3((unsigned)nwg), dim3(kKBlock), 0, ctx->stream, P);
        else hipLaunchKernelGGL(kmer_plain_kernel<kModeEmit>, dim3((unsigned)nwg), dim3(kKBlock), 0, ctx->stream, P);
        MP_HIP(hipGetLastError());
    }
    ctx->cur = 0;
    *n_out = n;
    return 0;
}

}  // namespace mp
