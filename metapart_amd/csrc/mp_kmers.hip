// Canonical k-mer kernels: m-mer histograms, digit counts and the ordered
// 3-minimizer enumeration.
//
// Plain variant (every k-mer; countMmers.c:83-217, enumerateKmers.c:127-371).
// The reference rolls a forward and a reverse-complement code through each
// read on one CPU thread.  Here the lanes of a workgroup split the flattened
// list of a group of reads' k-mer positions into equal contiguous runs, so
// the work per lane does not depend on read length (for_group_windows):
//   F  = 2K bits of the packed stream at the k-mer's first base (two 64-bit
//        loads and a funnel shift; MSB-first packing makes this the
//        reference's curKmerInt) for a lane's first window of a read, then
//        rolled one base at a time, as the reference does,
//   R  = reverse complement of F (bit reversal, then rolled: the reference's
//        curRcInt),
//   N  = bases since the last 'N' of the 1-bit 'N' stream must reach K (lastNPos),
//   canonical = F < R ? F : R, bucket = canonical >> 2(K-10) (getMmerPrefix).
// At K = 32 the reference's mask (1 << 64) - 1 evaluates to 0 on x86, so every
// forward code after the first position is 0; that is reproduced.
// The plain tuples themselves are not enumerated into memory: the first radix
// pass of the sort generates them from the reads in the same order (read
// order, then position order within a read).  This file gives it the digit
// counts: the m-mer histogram's counting pass also counts the k-mer bytes the
// sort reads, and kmer_digit_hist_kernel counts a scan that holds part of the
// buckets.
//
// 3-minimizer variant (countMmers_3minimizers.c:83-271,
// enumerateKmers_3minimizers.c:129-474): a sequential window state machine per
// read, one lane per read, over the same O(1) canonical k-mers; its tuples are
// enumerated (ordered stream compaction with offsets from a count pass).
#include "mp_internal.h"
#include "mp_kmer.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace mp {

// ---------------------------------------------------------------- m-mer histogram
// The per-chunk table has 2^20 bins per chunk, far beyond LDS, and one global
// atomic per k-mer is bound by the atomic rate of the memory side.  So the
// histogram is built in three LDS-privatised steps:
//   1. tiles of up to 1024 reads of one chunk count their k-mers per digit
//      (top 5 bits of the 20-bit bucket) in LDS,
//   2. a scan over (chunk, digit, tile) places every tile's digit run, and the
//      tiles write the low 15 bucket bits of each k-mer there (u16); few
//      digits keep few partially written lines open per tile,
//   3. each (chunk, digit) partition is counted in a 32K-bin LDS histogram and
//      written to its slice of the chunk's row.
// Bucket order inside a run does not matter to a histogram, so step 2 takes
// slots with LDS atomics instead of a stable ranking.
constexpr int kHDigitBits = 5;
constexpr int kHDigits = 1 << kHDigitBits;
constexpr int kHLowBits = 20 - kHDigitBits;
constexpr int kHLow = 1 << kHLowBits;
constexpr uint32_t kHReads = 1024;       // reads per tile (steps 1, 2)
constexpr uint64_t kHSlab = 4u << 20;    // items per workgroup (step 3)
constexpr int kHWords = 1024;            // staged stream words of a group of kRG reads (150 bp: ~450)

constexpr int kBhRep = 32;  // copies of the byte histograms (tile t adds to copy t % kBhRep)

struct HParams {
    unsigned long long *bh;     // kBhRep x 8 x 256 counts of the k-mer digits in bmask
    uint32_t bmask;             // bit 1 + b: count digit b (the radix passes of the scans)
    KeyLayout kl;               // the scans' key digits
    const uint64_t *chunk_ids;  // C+1 first read ids
    const uint32_t *tile0;      // C+1 first tile of each chunk
    int C;
    uint32_t *tcnt;             // [c][d][tile of c] digit counts
    const uint64_t *toff;       // exclusive scan of tcnt
    uint16_t *items;            // low bucket bits, grouped by (chunk, digit)
    // digit counts per tile of the generating radix pass (tiles of gtile
    // consecutive windows), for a scan over every k-mer: gcnt[tile][digit] of
    // the k-mer byte at gshift
    uint32_t *gcnt;
    uint64_t gtile;
    int gshift;
    // a batch: byte counts per chunk (bh holds C copies), and generating-pass
    // tiles over every chunk's own windows: chunk c's tiles start at gft[c],
    // its windows at gw[c] (C + 1 entries each)
    int per_chunk;
    const uint64_t *gft, *gw;
};

constexpr int kGenCntTiles = 24;  // generating-pass tiles a histogram tile counts in LDS (more: global atomics)

struct HPiece {
    uint64_t b, e;   // item range
    uint32_t part;   // chunk * kHDigits + digit
    uint32_t split;  // partition spread over several pieces: add atomically
};

// PHASE 0: digit counts per tile (GC: and the generating pass's digit counts
// per tile, H.gcnt, whose LDS copy would otherwise cost the plain form two
// thirds of its workgroups per CU), 1: scatter of the low bits
template <int PHASE, bool GC = false>
__global__ __launch_bounds__(kKBlock) void hist_tile_kernel(KParams P, HParams H)
{
    __shared__ uint32_t s_pref[kRG + 1];
    __shared__ uint64_t s_base[kRG];
    __shared__ uint32_t s_dig[kHDigits];
    __shared__ uint64_t s_out[kHDigits];
    __shared__ uint32_t s_bh[PHASE == 0 ? 8 : 1][kSortRadix];
    __shared__ uint64_t s_words[kHWords];  // the group's 2-bit words, then its 'N' words
    // digit d of a generating-pass tile in the 16-bit half d & 1 of word d / 2 (a
    // tile holds at most 2^16 - 1 windows): half the LDS, more workgroups per CU
    __shared__ uint32_t s_gc[GC ? kGenCntTiles : 1][kSortRadix / 2];
    static_assert(kSortRadix % 2 == 0, "two digits per word");
    const int tid = threadIdx.x;
    if (PHASE == 0) {
        for (int j = tid; j < 8 * kSortRadix; j += kKBlock) (&s_bh[0][0])[j] = 0;
        if (GC)
            for (int j = tid; j < kGenCntTiles * kSortRadix / 2; j += kKBlock) (&s_gc[0][0])[j] = 0;
    }
    const uint32_t t = blockIdx.x;
    int c = 0, hi = H.C - 1;  // chunk: last c with tile0[c] <= t
    while (c < hi) {
        const int mid = (c + hi + 1) >> 1;
        if (H.tile0[mid] <= t) c = mid; else hi = mid - 1;
    }
    const uint32_t tl = t - H.tile0[c], Tc = H.tile0[c + 1] - H.tile0[c];
    const uint64_t rb = H.chunk_ids[c] + (uint64_t)tl * kHReads;
    const uint64_t re = min(rb + kHReads, H.chunk_ids[c + 1]);
    const uint64_t mbase = (uint64_t)kHDigits * H.tile0[c] + tl;
    // the generating-pass tiles this tile's windows fall in: [g_lo, g_hi];
    // tiles count from `origin` (window) as tile `tbase`, up to window `cend`
    const uint64_t w_lo = P.pos_off[rb], w_hi = P.pos_off[re];
    const uint64_t origin = H.per_chunk ? H.gw[c] : 0, tbase = H.per_chunk ? H.gft[c] : 0;
    const uint64_t cend = H.per_chunk ? H.gw[c + 1] : ~0ull;
    const uint64_t g_lo = H.gcnt ? tbase + (w_lo - origin) / H.gtile : 0;
    if (tid < kHDigits) {
        s_dig[tid] = 0;
        if (PHASE == 1) s_out[tid] = H.toff[mbase + (uint64_t)tid * Tc];
    }
    for (uint64_t r0 = rb; r0 < re; r0 += kRG) {
        const int nr = (int)min((uint64_t)kRG, re - r0);
        __syncthreads();  // previous group done with s_pref / s_base
        // the group's read table and stream words (a group of very long reads
        // reads memory): every lane's loads are issued before its stores, the
        // indices clamped instead of branched on
        const uint64_t p0 = P.pos_off[r0];
        const uint64_t v_pos = P.pos_off[r0 + min(tid, nr)], v_base = P.base_off[r0 + min(tid, nr - 1)];
        const uint64_t b0 = P.base_off[r0], b1 = P.base_off[r0 + nr];
        const uint64_t wp0 = b0 >> 5, np = b1 > b0 ? ((b1 - 1) >> 5) + 2 - wp0 : 0;
        const uint64_t wn0 = b0 >> 6, nn = b1 > b0 ? ((b1 - 1) >> 6) + 2 - wn0 : 0;
        const bool staged = nn && np + nn <= (uint64_t)kHWords;
        constexpr int kHW = kHWords / kKBlock;
        static_assert(kHW * kKBlock == kHWords, "staged words per lane");
        uint64_t v_w[kHW];
        if (staged) {
#pragma unroll
            for (int q = 0; q < kHW; q++) {  // word j: 2-bit word j (< np), else 'N' word j - np
                const uint64_t j = (uint64_t)q * kKBlock + tid;
                v_w[q] = j < np ? P.packed[wp0 + j] : P.nmask[wn0 + min(j - np, nn - 1)];
            }
        }
        if (tid <= nr) s_pref[tid] = (uint32_t)(v_pos - p0);
        if (tid < nr) s_base[tid] = v_base;
        if (staged) {
#pragma unroll
            for (int q = 0; q < kHW; q++)
                if ((uint64_t)q * kKBlock + tid < np + nn) s_words[q * kKBlock + tid] = v_w[q];
        }
        __syncthreads();
        // a lane's windows are consecutive: its generating-pass tile advances at most once per gtile windows
        uint64_t gt = ~0ull, gnext = 0;
        auto body = [&](uint64_t canon, uint32_t q) {
            if (GC) {
                const uint64_t wi = p0 + q;
                if (gt == ~0ull) {
                    const uint64_t lt = (wi - origin) / H.gtile;
                    gt = tbase + lt;
                    gnext = origin + (lt + 1) * H.gtile;
                }
                while (wi >= gnext) {
                    gt++;
                    gnext += H.gtile;
                }
                const uint32_t d = (uint32_t)(canon >> H.gshift) & 0xFF;
                if (gt - g_lo < (uint64_t)kGenCntTiles) atomicAdd(&s_gc[gt - g_lo][d >> 1], 1u << (16 * (d & 1)));
                else atomicAdd(&H.gcnt[gt * kSortRadix + d], 1u);
            }
            const uint32_t b = (uint32_t)(canon >> P.bshift);
            const uint32_t d = b >> kHLowBits;
            const uint32_t slot = atomicAdd(&s_dig[d], 1u);
            if (PHASE == 1) H.items[s_out[d] + slot] = (uint16_t)(b & (kHLow - 1));
            if (PHASE == 0) {  // digit counts of a later sort over every k-mer
#pragma unroll
                for (int y = 0; y < 8; y++)
                    if ((H.bmask >> (y + 1)) & 1) atomicAdd(&s_bh[y][H.kl.digit(canon, y)], 1u);
            }
        };
        // the words from LDS or from memory in two copies of the loop: one
        // accessor choosing at run time reads both through flat loads, whose
        // waits also wait for the items' stores
        if (staged) {
            for_group_windows(
                P.K, s_pref, s_base, nr, tid, kKBlock, [&](uint64_t w) -> uint64_t { return s_words[np + (w - wn0)]; },
                [&](uint64_t w) -> uint64_t { return s_words[w - wp0]; }, body);
        } else {
            for_group_windows(
                P.K, s_pref, s_base, nr, tid, kKBlock, [&](uint64_t w) -> uint64_t { return P.nmask[w]; },
                [&](uint64_t w) -> uint64_t { return P.packed[w]; }, body);
        }
    }
    if (PHASE == 0) {
        __syncthreads();
        if (tid < kHDigits) H.tcnt[mbase + (uint64_t)tid * Tc] = s_dig[tid];
        unsigned long long *bh = H.bh + (uint64_t)(H.per_chunk ? c : t % kBhRep) * 8 * kSortRadix;
        for (int j = tid; j < 8 * kSortRadix; j += kKBlock) {
            const uint32_t v = (&s_bh[0][0])[j];
            if (v) atomicAdd(&bh[j], (unsigned long long)v);
        }
        if (GC && w_hi > w_lo) {
            // a tile whose windows all lie in this tile's is written; one shared
            // with a neighbour (the first and last) is added to
            const uint64_t g_hi = tbase + (w_hi - 1 - origin) / H.gtile;
            const uint64_t ng = min(g_hi - g_lo + 1, (uint64_t)kGenCntTiles);
            for (uint64_t j = tid; j < ng * kSortRadix; j += kKBlock) {
                const uint64_t g = g_lo + j / kSortRadix;
                const uint32_t dd = (uint32_t)(j % kSortRadix);
                const uint32_t v = (s_gc[j / kSortRadix][dd >> 1] >> (16 * (dd & 1))) & 0xFFFFu;
                const uint64_t gw0 = origin + (g - tbase) * H.gtile, gw1 = min(gw0 + H.gtile, cend);
                const bool owned = gw0 >= w_lo && gw1 <= w_hi;
                if (owned) H.gcnt[g * kSortRadix + (j % kSortRadix)] = v;
                else if (v) atomicAdd(&H.gcnt[g * kSortRadix + (j % kSortRadix)], v);
            }
        }
    }
}

// start of every (chunk, digit) partition, and the end of the last
__global__ void hist_bounds_kernel(HParams H, uint64_t *__restrict__ bounds)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > H.C * kHDigits) return;
    if (i == H.C * kHDigits) { bounds[i] = H.toff[(uint64_t)kHDigits * H.tile0[H.C]]; return; }
    const int c = i / kHDigits, d = i % kHDigits;
    const uint32_t Tc = H.tile0[c + 1] - H.tile0[c];
    bounds[i] = H.toff[(uint64_t)kHDigits * H.tile0[c] + (uint64_t)d * Tc];
}

__global__ __launch_bounds__(1024) void hist_count_kernel(const uint16_t *__restrict__ items,
                                                          const HPiece *__restrict__ pieces,
                                                          uint32_t *__restrict__ counts)
{
    __shared__ uint32_t s_h[kHLow];
    const HPiece pc = pieces[blockIdx.x];
    for (int i = threadIdx.x; i < kHLow; i += blockDim.x) s_h[i] = 0;
    __syncthreads();
    // 8 items per 16-byte load on the aligned middle, single items at the ends
    const uint64_t a = min(pc.e, (pc.b + 7) & ~7ull), z = max(a, pc.e & ~7ull);
    for (uint64_t i = pc.b + threadIdx.x; i < a; i += blockDim.x) atomicAdd(&s_h[items[i]], 1u);
    for (uint64_t i = z + threadIdx.x; i < pc.e; i += blockDim.x) atomicAdd(&s_h[items[i]], 1u);
    const uint4 *v = (const uint4 *)(items + a);
    const uint64_t nv = (z - a) / 8;
    for (uint64_t j = threadIdx.x; j < nv; j += blockDim.x) {
        const uint4 w = v[j];
        const uint32_t x[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            atomicAdd(&s_h[x[k] & 0xFFFF], 1u);
            atomicAdd(&s_h[x[k] >> 16], 1u);
        }
    }
    __syncthreads();
    const uint32_t c = pc.part / kHDigits, d = pc.part % kHDigits;
    uint32_t *row = counts + (size_t)c * (MP_NBUCKETS + 2) + ((size_t)d << kHLowBits);
    for (int i = threadIdx.x; i < kHLow; i += blockDim.x) {
        const uint32_t h = s_h[i];
        if (!pc.split) row[i] = h;
        else if (h) atomicAdd(&row[i], h);
    }
}

// seg: a batch (chunk = sample): only the first phase runs, for the byte
// counts per sample (khist, C copies) and the generating pass's digit counts
// over tiles of every sample's own windows (d_counts unused)
__global__ void gather_u64_kernel(const uint64_t *__restrict__ src, const uint64_t *__restrict__ idx, uint64_t n,
                                  uint64_t *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = src[idx[i]];
}

static int count_plain_partitioned(mp_ctx *ctx, KParams P, const uint64_t *chunk_first_id, int C, uint32_t *d_counts,
                                   bool seg = false)
{
    std::vector<uint32_t> tile0(C + 1);
    tile0[0] = 0;
    for (int c = 0; c < C; c++) {
        const uint64_t nr = chunk_first_id[c + 1] - chunk_first_id[c];
        tile0[c + 1] = tile0[c] + (uint32_t)((nr + kHReads - 1) / kHReads);
    }
    const uint64_t NT = tile0[C];
    if (NT == 0) return 0;
    const uint64_t M = (uint64_t)kHDigits * NT;
    const uint64_t nparts = (uint64_t)C * kHDigits;
    // hpart: chunk ids (C+1 u64) | tile0 (C+1 u32, padded) | bounds (nparts+1 u64) | tile counts (M u32)
    const uint64_t o_ids = 0, o_t0 = o_ids + 8 * (C + 1), o_bd = o_t0 + 8 * ((C + 2) / 2),
                   o_tc = o_bd + 8 * (nparts + 1), bytes = o_tc + 4 * M + 64;
    MP_TRY(ensure(ctx, ctx->hpart, bytes));
    MP_TRY(ensure(ctx, ctx->wg_off, (M + 1) * sizeof(uint64_t)));
    uint8_t *hb = (uint8_t *)ctx->hpart.p;
    HParams H{};
    const uint64_t ncopies = seg ? (uint64_t)C : (uint64_t)kBhRep;
    MP_TRY(ensure(ctx, ctx->khist, sizeof(uint64_t) * ncopies * 8 * kSortRadix));
    MP_TRY(fill_u32(ctx, ctx->khist.p, 0, 2 * ncopies * 8 * kSortRadix));
    H.bh = (unsigned long long *)ctx->khist.p;
    H.bmask = scan_pass_mask(ctx->cfg.k) & 0x1FEu;
    H.kl = scan_layout(ctx);
    H.per_chunk = seg ? 1 : 0;
    ctx->khist_copies = (int)ncopies;
    // digit counts per generating-pass tile of the lowest sorted byte, for a
    // batch sorted sample by sample (its generating pass then takes counted
    // offsets).  A plain run does not count them: at 2.4e9 windows the
    // counted generating pass measured equal to the look-back one (27.9 ms),
    // and counting them cost the histogram about 4 ms (LDS for 24 tiles).
    ctx->gcnt_valid = false;
    ctx->gcnt_seg = seg;
    if (H.bmask && !ctx->arena && seg) {
        const uint64_t gtile = sort_tile_items(ctx);
        if (gtile >= (1u << 16)) return ctx->fail(MP_EINVAL, "radix tile of %llu windows: counts need 16 bits",
                                                 (unsigned long long)gtile);
        uint64_t npos = 0;
        MP_HIP(hipMemcpyAsync(&npos, (const uint64_t *)ctx->pos_off.p + ctx->N, sizeof(uint64_t),
                              hipMemcpyDeviceToHost, ctx->stream));
        MP_HIP(hipStreamSynchronize(ctx->stream));
        uint64_t gtiles = (npos + gtile - 1) / gtile;
        if (seg) {  // every sample's windows in tiles of their own
            ctx->seg_wpos.assign(C + 1, 0);
            MP_TRY(ensure(ctx, ctx->segtab, 3 * (uint64_t)(C + 1) * sizeof(uint64_t)));
            uint64_t *d_w = (uint64_t *)ctx->segtab.p, *d_ids = d_w + 2 * (C + 1);
            MP_HIP(hipMemcpyAsync(d_ids, chunk_first_id, (C + 1) * sizeof(uint64_t), hipMemcpyHostToDevice,
                                  ctx->stream));
            hipLaunchKernelGGL(gather_u64_kernel, dim3((C + 256) / 256), dim3(256), 0, ctx->stream,
                               (const uint64_t *)ctx->pos_off.p, (const uint64_t *)d_ids, (uint64_t)(C + 1),
                               d_w + (C + 1));
            MP_HIP(hipGetLastError());
            MP_HIP(hipMemcpyAsync(ctx->seg_wpos.data(), d_w + (C + 1), (C + 1) * sizeof(uint64_t),
                                  hipMemcpyDeviceToHost, ctx->stream));
            MP_HIP(hipStreamSynchronize(ctx->stream));
            std::vector<uint64_t> gft(C + 1, 0);
            for (int c = 0; c < C; c++)
                gft[c + 1] = gft[c] + (ctx->seg_wpos[c + 1] - ctx->seg_wpos[c] + gtile - 1) / gtile;
            gtiles = gft[C];
            MP_HIP(hipMemcpyAsync(d_w, gft.data(), (C + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
            H.gft = d_w;
            H.gw = d_w + (C + 1);
            MP_HIP(hipStreamSynchronize(ctx->stream));  // gft is read by the copy above
        }
        if (gtiles) {
            MP_TRY(ensure(ctx, ctx->gcnt, gtiles * kSortRadix * sizeof(uint32_t)));
            MP_TRY(fill_u32(ctx, ctx->gcnt.p, 0, gtiles * kSortRadix));
            H.gcnt = (uint32_t *)ctx->gcnt.p;
            H.gtile = gtile;
            H.gshift = H.kl.shift[__builtin_ctz(H.bmask) - 1];  // the generating pass's digit: a whole byte
            ctx->gcnt_tile = gtile;
            ctx->gcnt_tiles = gtiles;
            ctx->gcnt_byte = __builtin_ctz(H.bmask) - 1;
        }
    }
    H.chunk_ids = (const uint64_t *)(hb + o_ids);
    H.tile0 = (const uint32_t *)(hb + o_t0);
    H.C = C;
    H.tcnt = (uint32_t *)(hb + o_tc);
    H.toff = (const uint64_t *)ctx->wg_off.p;
    uint64_t *d_bounds = (uint64_t *)(hb + o_bd);
    MP_HIP(hipMemcpyAsync(hb + o_ids, chunk_first_id, 8 * (C + 1), hipMemcpyHostToDevice, ctx->stream));
    MP_HIP(hipMemcpyAsync(hb + o_t0, tile0.data(), 4 * (C + 1), hipMemcpyHostToDevice, ctx->stream));
    if (H.gcnt) hipLaunchKernelGGL((hist_tile_kernel<0, true>), dim3((unsigned)NT), dim3(kKBlock), 0, ctx->stream, P, H);
    else hipLaunchKernelGGL((hist_tile_kernel<0, false>), dim3((unsigned)NT), dim3(kKBlock), 0, ctx->stream, P, H);
    MP_HIP(hipGetLastError());
    MP_TRY(scan_u32(ctx, H.tcnt, (uint64_t *)ctx->wg_off.p, M));
    hipLaunchKernelGGL(hist_bounds_kernel, dim3((unsigned)((nparts + 256) / 256)), dim3(256), 0, ctx->stream, H,
                       d_bounds);
    MP_HIP(hipGetLastError());
    std::vector<uint64_t> bounds(nparts + 1);
    MP_HIP(hipMemcpyAsync(bounds.data(), d_bounds, 8 * (nparts + 1), hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    const uint64_t total = bounds[nparts];
    ctx->khist_mask = H.bmask;
    ctx->khist_valid = true;
    ctx->gcnt_valid = H.gcnt != nullptr;
    if (seg) return 0;  // a batch's counts: nothing more
    ctx->kmer_total = total;
    if (total == 0) return 0;
    // the low bits stage in the second tuple buffer, unused until the first scan's sort
    MP_TRY(ensure(ctx, ctx->keys[1], total * sizeof(uint16_t) + 64));
    H.items = (uint16_t *)ctx->keys[1].p;
    hipLaunchKernelGGL(hist_tile_kernel<1>, dim3((unsigned)NT), dim3(kKBlock), 0, ctx->stream, P, H);
    MP_HIP(hipGetLastError());
    std::vector<HPiece> pieces;
    for (uint64_t p = 0; p < nparts; p++) {
        const uint64_t b = bounds[p], e = bounds[p + 1];
        if (e == b) continue;
        const uint64_t n = (e - b + kHSlab - 1) / kHSlab;
        for (uint64_t j = 0; j < n; j++)
            pieces.push_back({b + j * kHSlab, std::min(e, b + (j + 1) * kHSlab), (uint32_t)p, n > 1 ? 1u : 0u});
    }
    HPiece *d_pieces = nullptr;
    MP_TRY(tmp_alloc(ctx, (void **)&d_pieces, pieces.size() * sizeof(HPiece)));
    MP_HIP(hipMemcpyAsync(d_pieces, pieces.data(), pieces.size() * sizeof(HPiece), hipMemcpyHostToDevice,
                          ctx->stream));
    hipLaunchKernelGGL(hist_count_kernel, dim3((unsigned)pieces.size()), dim3(1024), 0, ctx->stream,
                       (const uint16_t *)H.items, (const HPiece *)d_pieces, d_counts);
    MP_HIP(hipGetLastError());
    MP_TRY(tmp_free(ctx, d_pieces));
    // the host vectors above must outlive the copies
    MP_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

// ---------------------------------------------------------------- 3-minimizer
// The reference rolls forward and reverse codes through the read
// (countMmers_3minimizers.c:107-190); so does the lane of a read here
// (RollKmer over the packed streams, 8x smaller than the FASTQ bytes, so a
// lane walking its read stays in cache).  The window state machine itself is
// sequential per read and is restated as is.

// counter: countMmers_3minimizers.c:83-271 (no range filter)
template <typename Src, typename Sink>
__device__ void min3_count_read(int L, int K, int bshift, Src kmer, Sink sink)
{
    if (L <= K) return;
    const int last = L - K;
    uint64_t minK = 0, minB = 0, c;
    bool fwd;
    int next = last / 3;
    if (kmer(0, c, fwd)) {
        minK = c;
        minB = c >> bshift;
    }
    for (int i = 1; i <= last; i++) {
        if (!kmer(i, c, fwd)) continue;  // an 'N' in the window (i <= lastN)
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
    if (minK != 0) sink(minB);
}

// enumerator: enumerateKmers_3minimizers.c:129-474 (range filter before the
// window logic).  The reference keeps the minimum's neighbour characters,
// complemented when the minimum is a reverse k-mer; here the minimum's
// position and strand are kept and the characters read when it is emitted.
// An emission inside the loop orders the pair by the strand of the k-mer that
// closes the window, the final one always as (next, prev), as the reference.
template <typename Src, typename Sink>
__device__ void min3_enum_read(const uint8_t *s, int L, int K, int bshift, uint64_t lo, uint64_t hi, Src kmer,
                               Sink sink)
{
    if (L <= K) return;
    const int last = L - K;
    uint64_t minK = 0, c;
    int minPos = 0;
    bool minFwd = true, fwd;
    auto neighbours = [&](uint8_t &cp, uint8_t &cn) {
        const uint8_t p = minPos == 0 ? (uint8_t)'N' : s[minPos - 1];
        const uint8_t n = minPos == last ? (uint8_t)'N' : s[minPos + K];
        cp = minFwd ? p : rc_char(p);
        cn = minFwd ? n : rc_char(n);
    };
    if (kmer(0, c, fwd)) {
        const uint64_t b = c >> bshift;
        if (b >= lo && b < hi) {
            minK = c;
            minPos = 0;
            minFwd = fwd;
        }
    }
    int nextW = last / 3;
    for (int i = 1; i <= last; i++) {
        if (!kmer(i, c, fwd)) continue;
        const uint64_t b = c >> bshift;
        if (!(b >= lo && b < hi)) continue;
        if (i > nextW) {
            if (minK != 0) {
                uint8_t cp, cn;
                neighbours(cp, cn);
                sink(minK, fwd ? adj_of(cp, cn) : adj_of(cn, cp));
            }
            minK = c;
            minPos = i;
            minFwd = fwd;
            nextW += last / 3;
            if (nextW + last / 3 > last) nextW = last;
        } else if (c < minK || minK == 0) {
            minK = c;
            minPos = i;
            minFwd = fwd;
        }
    }
    if (minK != 0) {
        uint8_t cp, cn;
        neighbours(cp, cn);
        sink(minK, adj_of(cn, cp));
    }
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
    uint64_t b0 = 0;
    if (r < P.N) {
        s = P.fq + P.seq_off[r];
        L = (int)P.seq_len[r];
        b0 = P.base_off[r];
    }
    // the window logic asks for positions 0, 1, 2, ... in order: roll the k-mer
    RollKmer rk;
    int cur = -1;
    auto nword = [&](uint64_t w) { return P.nmask[w]; };
    auto pword = [&](uint64_t w) { return P.packed[w]; };
    auto kmer = [&](int i, uint64_t &c, bool &fwd) {
        if (cur < 0 || i < cur) {
            rk.init(P.K, b0 + (uint64_t)i, nword, pword);
            cur = i;
        }
        while (cur < i) {
            cur++;
            rk.roll(P.K, b0 + (uint64_t)cur + P.K - 1, nword, pword);
        }
        if (!rk.valid(P.K)) return false;
        const uint64_t f = (P.K == 32 && i > 0) ? 0 : rk.F;
        fwd = f < rk.R;
        c = fwd ? f : rk.R;
        return true;
    };
    // the chunk of read r (chunk_ids: C + 1 first read ids)
    auto chunk_of = [&](uint64_t rr) {
        int lo = 0, hi = P.C - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (P.chunk_ids[mid] <= rr) lo = mid; else hi = mid - 1;
        }
        return lo;
    };
    if (MODE == kModeHist) {
        if (r < P.N) {
            uint32_t *cnts = P.counts + (size_t)chunk_of(r) * (MP_NBUCKETS + 2);
            min3_count_read(L, P.K, P.bshift, kmer, [&](uint64_t b) { atomicAdd(&cnts[b], 1u); });
        }
        return;
    }
    if (MODE == kModeTop) {  // counts[c]: 1 + the highest bucket of chunk c's minimizers
        if (r < P.N) {
            uint32_t *top = P.counts + chunk_of(r);
            min3_count_read(L, P.K, P.bshift, kmer, [&](uint64_t b) { atomicMax(top, (uint32_t)b + 1u); });
        }
        return;
    }
    // the scan's bucket range, or the read's chunk's own (a batch of samples)
    const uint64_t rhi = P.chunk_hi && r < P.N ? P.chunk_hi[chunk_of(r)] : P.hi;
    if (r < P.N) min3_enum_read(s, L, P.K, P.bshift, P.lo, rhi, kmer, [&](uint64_t, uint8_t) { cnt++; });
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
    const uint32_t label = label_of(P, r);
    min3_enum_read(s, L, P.K, P.bshift, P.lo, rhi, kmer, [&](uint64_t k, uint8_t a) {
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

int batch_counts(mp_ctx *ctx)
{
    KParams P = base_params(ctx);
    P.C = ctx->C;
    return count_plain_partitioned(ctx, P, ctx->chunk_ids.data(), ctx->C, nullptr, true);
}

int count_chunks_device(mp_ctx *ctx, const uint64_t *chunk_first_id, int C, uint32_t *d_counts)
{
    KParams P = base_params(ctx);
    uint64_t *d_ids = nullptr;
    MP_TRY(tmp_alloc(ctx, (void **)&d_ids, sizeof(uint64_t) * (C + 1)));
    MP_HIP(hipMemcpyAsync(d_ids, chunk_first_id, sizeof(uint64_t) * (C + 1), hipMemcpyHostToDevice, ctx->stream));
    MP_TRY(fill_u32(ctx, d_counts, 0, (uint64_t)C * (MP_NBUCKETS + 2)));
    P.counts = d_counts;
    P.chunk_ids = d_ids;
    P.C = C;
    if (ctx->cfg.min3) {
        hipLaunchKernelGGL(kmer_min3_kernel<kModeHist>, dim3((unsigned)((ctx->N + kKBlock - 1) / kKBlock)),
                           dim3(kKBlock), 0, ctx->stream, P);
        MP_HIP(hipGetLastError());
    } else {
        MP_TRY(count_plain_partitioned(ctx, P, chunk_first_id, C, d_counts));
    }
    MP_TRY(tmp_free(ctx, d_ids));
    return 0;
}

// Digit counts of a scan's tuples for the radix sort whose first pass
// generates them: every in-range k-mer of every read, counted per k-mer byte
// of the pass mask in LDS (groups of kRG reads, grid-stride), then added to
// ghist[pass][digit] (pass 1 + b for byte b).
__global__ __launch_bounds__(kKBlock) void kmer_digit_hist_kernel(KParams P, uint32_t pass_mask, KeyLayout kl,
                                                                  unsigned long long *__restrict__ ghist)
{
    __shared__ uint32_t h[8][kSortRadix];
    __shared__ uint32_t s_pref[kRG + 1];
    __shared__ uint64_t s_base[kRG];
    __shared__ uint64_t s_words[kHWords];  // the group's 2-bit words, then its 'N' words
    const int tid = threadIdx.x;
    for (int j = tid; j < 8 * kSortRadix; j += kKBlock) (&h[0][0])[j] = 0;
    const uint64_t ngrp = (P.N + kRG - 1) / kRG;
    for (uint64_t g = blockIdx.x; g < ngrp; g += gridDim.x) {
        const uint64_t r0 = g * kRG;
        const int nr = (int)((P.N - r0) < (uint64_t)kRG ? (P.N - r0) : (uint64_t)kRG);
        __syncthreads();  // the previous group is done with s_pref / s_base
        const uint64_t p0 = P.pos_off[r0];
        if (tid <= nr) s_pref[tid] = (uint32_t)(P.pos_off[r0 + tid] - p0);
        if (tid < nr) s_base[tid] = P.base_off[r0 + tid];
        const uint64_t b0 = P.base_off[r0], b1 = P.base_off[r0 + nr];
        const uint64_t wp0 = b0 >> 5, np = b1 > b0 ? ((b1 - 1) >> 5) + 2 - wp0 : 0;
        const uint64_t wn0 = b0 >> 6, nn = b1 > b0 ? ((b1 - 1) >> 6) + 2 - wn0 : 0;
        const bool staged = np + nn <= (uint64_t)kHWords;
        if (staged) {
            for (uint32_t j = tid; j < np; j += kKBlock) s_words[j] = P.packed[wp0 + j];
            for (uint32_t j = tid; j < nn; j += kKBlock) s_words[np + j] = P.nmask[wn0 + j];
        }
        __syncthreads();
        auto body = [&](uint64_t canon, uint32_t) {
            const uint64_t bk = canon >> P.bshift;
            if (bk < P.lo || bk >= P.hi) return;
#pragma unroll
            for (int b = 0; b < 8; b++)
                if ((pass_mask >> (b + 1)) & 1) atomicAdd(&h[b][kl.digit(canon, b)], 1u);
        };
        if (staged) {  // two copies of the loop, as in hist_tile_kernel
            for_group_windows(
                P.K, s_pref, s_base, nr, tid, kKBlock, [&](uint64_t w) -> uint64_t { return s_words[np + (w - wn0)]; },
                [&](uint64_t w) -> uint64_t { return s_words[w - wp0]; }, body);
        } else {
            for_group_windows(
                P.K, s_pref, s_base, nr, tid, kKBlock, [&](uint64_t w) -> uint64_t { return P.nmask[w]; },
                [&](uint64_t w) -> uint64_t { return P.packed[w]; }, body);
        }
    }
    __syncthreads();
    for (int j = tid; j < 8 * kSortRadix; j += kKBlock) {
        const uint32_t v = (&h[0][0])[j];
        if (v) atomicAdd(&ghist[kSortRadix + j], (unsigned long long)v);
    }
}

// Per-bucket counts of the k-mers in [lo, hi) by device atomics (for a scan
// that must be split into parts; no staging memory).
__global__ __launch_bounds__(kKBlock) void kmer_bucket_count_kernel(KParams P, unsigned long long *__restrict__ cnt)
{
    __shared__ uint32_t s_pref[kRG + 1];
    __shared__ uint64_t s_base[kRG];
    const int tid = threadIdx.x;
    const uint64_t ngrp = (P.N + kRG - 1) / kRG;
    for (uint64_t g = blockIdx.x; g < ngrp; g += gridDim.x) {
        const uint64_t r0 = g * kRG;
        const int nr = (int)((P.N - r0) < (uint64_t)kRG ? (P.N - r0) : (uint64_t)kRG);
        __syncthreads();  // the previous group is done with s_pref / s_base
        const uint64_t p0 = P.pos_off[r0];
        if (tid <= nr) s_pref[tid] = (uint32_t)(P.pos_off[r0 + tid] - p0);
        if (tid < nr) s_base[tid] = P.base_off[r0 + tid];
        __syncthreads();
        for_group_windows(
            P.K, s_pref, s_base, nr, tid, kKBlock, [&](uint64_t w) { return P.nmask[w]; },
            [&](uint64_t w) { return P.packed[w]; },
            [&](uint64_t canon, uint32_t) {
                const uint64_t bk = canon >> P.bshift;
                if (bk >= P.lo && bk < P.hi) atomicAdd(&cnt[bk - P.lo], 1ull);
            });
    }
}

__global__ void label_snap_kernel(const uint32_t *__restrict__ parent, const uint8_t *__restrict__ bad, uint64_t N,
                                  int scan0, uint32_t *__restrict__ out)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < N) out[r] = scan0 ? (uint32_t)r : (bad[r] ? (uint32_t)r : parent[r]);
}

int snapshot_labels(mp_ctx *ctx, int s)
{
    MP_TRY(ensure(ctx, ctx->labsnap, ctx->N * sizeof(uint32_t) + 64));
    hipLaunchKernelGGL(label_snap_kernel, dim3((unsigned)((ctx->N + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const uint32_t *)ctx->parent.p, (const uint8_t *)ctx->bad.p, ctx->N, s == 0 ? 1 : 0,
                       (uint32_t *)ctx->labsnap.p);
    MP_HIP(hipGetLastError());
    ctx->use_labsnap = true;
    return 0;
}

int bucket_prefix_range(mp_ctx *ctx, uint64_t lo, uint64_t hi, std::vector<uint64_t> &pre)
{
    const uint64_t nb = hi - lo;
    pre.assign(nb + 1, 0);
    if (ctx->cfg.min3) {  // the 3-minimizer counter adds per chunk with atomics already
        const int C = ctx->C;
        const uint64_t words = (uint64_t)C * (MP_NBUCKETS + 2);
        MP_TRY(ensure(ctx, ctx->scratch, words * sizeof(uint32_t) + 4096));
        MP_TRY(ensure(ctx, ctx->hist, (2 * (uint64_t)MP_NBUCKETS + 2) * sizeof(uint64_t)));
        uint32_t *d_counts = (uint32_t *)ctx->scratch.p;
        MP_TRY(count_chunks_device(ctx, ctx->chunk_ids.data(), C, d_counts));
        uint64_t *d_tot = (uint64_t *)ctx->hist.p;
        MP_TRY(bucket_totals(ctx, d_counts, C, d_tot));
        std::vector<uint64_t> tot(MP_NBUCKETS);
        MP_HIP(hipMemcpyAsync(tot.data(), d_tot, sizeof(uint64_t) * MP_NBUCKETS, hipMemcpyDeviceToHost, ctx->stream));
        MP_HIP(hipStreamSynchronize(ctx->stream));
        for (uint64_t b = 0; b < nb; b++) pre[b + 1] = pre[b] + (lo + b < MP_NBUCKETS ? tot[lo + b] : 0);
        return 0;
    }
    KParams P = base_params(ctx);
    P.lo = lo;
    P.hi = hi;
    MP_TRY(ensure(ctx, ctx->hist, (nb + 64) * sizeof(uint64_t)));
    unsigned long long *d_cnt = (unsigned long long *)ctx->hist.p;
    MP_TRY(fill_u32(ctx, d_cnt, 0, 2 * nb));
    const uint64_t ngrp = (ctx->N + kRG - 1) / kRG;
    const unsigned blocks = (unsigned)(ngrp < 8192 ? ngrp : 8192);
    if (blocks) {
        hipLaunchKernelGGL(kmer_bucket_count_kernel, dim3(blocks), dim3(kKBlock), 0, ctx->stream, P, d_cnt);
        MP_HIP(hipGetLastError());
    }
    std::vector<uint64_t> c(nb);
    MP_HIP(hipMemcpyAsync(c.data(), d_cnt, sizeof(uint64_t) * nb, hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    for (uint64_t b = 0; b < nb; b++) pre[b + 1] = pre[b] + c[b];
    return 0;
}

int gen_histograms(mp_ctx *ctx, int s, uint64_t lo, uint64_t hi, uint32_t pass_mask, bool all, GenSource *g,
                   uint64_t *n_out)
{
    if (!pass_mask || (pass_mask & ~0x1FEu)) return ctx->fail(MP_EINVAL, "generated sorts take k-mer byte passes only");
    KParams P = base_params(ctx);
    P.lo = lo;
    P.hi = hi;
    P.scan0 = s == 0;
    P.parent = (const uint32_t *)ctx->parent.p;
    P.bad = (const uint8_t *)ctx->bad.p;
    P.labels = ctx->use_labsnap ? (const uint32_t *)ctx->labsnap.p : nullptr;
    MP_HIP(hipMemcpyAsync(&g->npos, (const uint64_t *)ctx->pos_off.p + ctx->N, sizeof(uint64_t),
                          hipMemcpyDeviceToHost, ctx->stream));
    if (all && ctx->khist_valid && !(pass_mask & ~ctx->khist_mask)) {
        // every k-mer is in the scan: the m-mer histogram counted these digits already
        const int nc = ctx->khist_copies;
        std::vector<uint64_t> rep((size_t)nc * 8 * kSortRadix);
        MP_HIP(hipMemcpyAsync(rep.data(), ctx->khist.p, rep.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
        MP_HIP(hipStreamSynchronize(ctx->stream));
        memset(g->h, 0, sizeof(g->h));
        for (int c = 0; c < nc; c++)
            for (int j = 0; j < 8 * kSortRadix; j++) g->h[kSortRadix + j] += rep[(size_t)c * 8 * kSortRadix + j];
        if (ctx->gcnt_seg) {  // per sample (chunk): tuples of each, from its byte counts
            const int b0 = __builtin_ctz(pass_mask);
            g->seg_tuples.assign(nc, 0);
            for (int c = 0; c < nc; c++)
                for (int r = 0; r < kSortRadix; r++)
                    g->seg_tuples[c] += rep[(size_t)c * 8 * kSortRadix + (size_t)(b0 - 1) * kSortRadix + r];
        }
        if (ctx->gcnt_valid && ctx->gcnt_tile == sort_tile_items(ctx)) {  // and per generating-pass tile
            g->gcnt = (const uint32_t *)ctx->gcnt.p;
            g->gcnt_tiles = ctx->gcnt_tiles;
            g->gcnt_pass = 1 + ctx->gcnt_byte;
            g->gcnt_seg = ctx->gcnt_seg;
        }
    } else {
        MP_TRY(ensure(ctx, ctx->hist, sizeof(uint64_t) * (kSortPasses * kSortRadix + 64)));
        unsigned long long *gh = (unsigned long long *)ctx->hist.p;
        MP_TRY(fill_u32(ctx, gh, 0, 2 * kSortPasses * kSortRadix));
        const uint64_t ngrp = (ctx->N + kRG - 1) / kRG;
        const unsigned blocks = (unsigned)(ngrp < 4096 ? ngrp : 4096);
        if (blocks) {
            hipLaunchKernelGGL(kmer_digit_hist_kernel, dim3(blocks), dim3(kKBlock), 0, ctx->stream, P, pass_mask,
                               scan_layout(ctx), gh);
            MP_HIP(hipGetLastError());
        }
        MP_HIP(hipMemcpyAsync(g->h, gh, sizeof(g->h), hipMemcpyDeviceToHost, ctx->stream));
        MP_HIP(hipStreamSynchronize(ctx->stream));
    }
    uint64_t n = 0;
    const int b0 = __builtin_ctz(pass_mask);
    for (int r = 0; r < kSortRadix; r++) n += g->h[b0 * kSortRadix + r];
    for (int b = 0; b < 2; b++) {
        MP_TRY(ensure(ctx, ctx->keys[b], (n + 1) * sizeof(uint64_t)));
        MP_TRY(ensure(ctx, ctx->lab[b], (n + 1) * sizeof(uint32_t)));
        MP_TRY(ensure(ctx, ctx->adj[b], (n + 16) * sizeof(uint8_t)));
    }
    g->P = P;
    *n_out = n;
    return 0;
}

static int enumerate_tuples(mp_ctx *ctx, int s, KParams &P, uint64_t *n_out);

int enumerate_scan(mp_ctx *ctx, int s, uint64_t lo, uint64_t hi, uint64_t *n_out, bool per_chunk_hi)
{
    if (!ctx->cfg.min3) return ctx->fail(MP_EINVAL, "k-mer tuples of every window come from the generating sort pass");
    KParams P = base_params(ctx);
    P.lo = lo;
    P.hi = hi;
    uint64_t *d_ids = nullptr;
    uint32_t *d_top = nullptr;
    int rc = 0;
    if (per_chunk_hi) {
        // A batch: the window logic of enumerateKmers_3minimizers.c skips the
        // k-mers outside the scan's bucket range, so every sample takes the
        // range of its own one-scan plan, [0, 1 + its highest minimizer
        // bucket) (mp_plan_ranges over its histogram), not the whole table
        const int C = ctx->C;
        MP_TRY(tmp_alloc(ctx, (void **)&d_ids, sizeof(uint64_t) * (C + 1) + sizeof(uint32_t) * (C + 1)));
        d_top = (uint32_t *)(d_ids + (C + 1));
        auto ranges = [&]() -> int {
            MP_HIP(hipMemcpyAsync(d_ids, ctx->chunk_ids.data(), sizeof(uint64_t) * (C + 1), hipMemcpyHostToDevice,
                                  ctx->stream));
            MP_TRY(fill_u32(ctx, d_top, 0, C + 1));
            P.chunk_ids = d_ids;
            P.C = C;
            P.counts = d_top;
            hipLaunchKernelGGL(kmer_min3_kernel<kModeTop>, dim3((unsigned)((ctx->N + kKBlock - 1) / kKBlock)),
                               dim3(kKBlock), 0, ctx->stream, P);
            MP_HIP(hipGetLastError());
            P.counts = nullptr;
            P.chunk_hi = d_top;
            return 0;
        };
        rc = ranges();
    }
    if (rc == 0) rc = enumerate_tuples(ctx, s, P, n_out);
    const int rf = tmp_free(ctx, d_ids);  // on the error paths too
    return rc ? rc : rf;
}

// the count, scan and emit launches of enumerate_scan over the prepared parameters
static int enumerate_tuples(mp_ctx *ctx, int s, KParams &P, uint64_t *n_out)
{
    P.scan0 = s == 0;
    P.parent = (const uint32_t *)ctx->parent.p;
    P.bad = (const uint8_t *)ctx->bad.p;
    P.labels = ctx->use_labsnap ? (const uint32_t *)ctx->labsnap.p : nullptr;
    const uint64_t nwg = (ctx->N + kKBlock - 1) / kKBlock;
    MP_TRY(ensure(ctx, ctx->wg_cnt, nwg * sizeof(uint32_t)));
    MP_TRY(ensure(ctx, ctx->wg_off, (nwg + 1) * sizeof(uint64_t)));
    P.wg_cnt = (uint32_t *)ctx->wg_cnt.p;
    P.wg_off = (const uint64_t *)ctx->wg_off.p;
    hipLaunchKernelGGL(kmer_min3_kernel<kModeCount>, dim3((unsigned)nwg), dim3(kKBlock), 0, ctx->stream, P);
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
        hipLaunchKernelGGL(kmer_min3_kernel<kModeEmit>, dim3((unsigned)nwg), dim3(kKBlock), 0, ctx->stream, P);
        MP_HIP(hipGetLastError());
    }
    ctx->cur = 0;
    *n_out = n;
    return 0;
}

}  // namespace mp
