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
//     never cleared between passes).  With ~500 tiles in flight a walk
//     passes tens of tiles that have only their own count published, and
//     each step is a device-coherent load from another CU (often another
//     XCD), so the walk reads 4 predecessors per step with their loads in
//     flight together (measured at 2.4e9 tuples: 21.5 ms per pass against
//     23.0 for one, 23.5 for 8 and 48.9 for 16 -- wider windows read status
//     words past the inclusive one, and that traffic competes with the
//     tuples'),
//   - tuples are reordered by digit in LDS and written out, so consecutive
//     lanes write consecutive addresses of each digit's run.
// Passes whose digit is constant over the scan (e.g. the top byte when K<29)
// are skipped: a stable pass on a constant digit is the identity.
//
// Batches of samples add a last pass whose digit is the sample of the tuple's
// label (labels are read ids, samples are contiguous id ranges): the result is
// every sample's tuples in (k-mer, adj) order, samples one after another, the
// same as sorting each sample alone.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "mp_internal.h"
#include "mp_kmer.h"

namespace mp {

constexpr int kRadix = kSortRadix;
constexpr int kPasses = kSortPasses;  // adj, k-mer bytes 0..7, sample of the label

constexpr uint64_t kStAgg = 1, kStInc = 2;
#ifndef MP_LOOKBACK
#define MP_LOOKBACK 4
#endif
constexpr int kLookBack = MP_LOOKBACK;  // predecessor tiles read per look-back step
// Look-back status words: epoch | flag (none, aggregate, inclusive) | value.
// 64-bit: 16-bit epoch, 46-bit value.  32-bit: 4-bit epoch, 26-bit value,
// half the look-back traffic; used when no digit of the pass has 2^26 tuples.
template <bool S32> struct Status;
template <> struct Status<false> {
    using T = uint64_t;
    __device__ static T pack(uint32_t epoch, uint64_t flag, uint64_t v) { return ((uint64_t)epoch << 48) | (flag << 46) | v; }
    __device__ static uint32_t flag(T s, uint32_t epoch)
    {
        return (uint32_t)(s >> 48) == epoch ? (uint32_t)((s >> 46) & 3) : 0u;
    }
    __device__ static uint64_t val(T s) { return s & ((1ull << 46) - 1); }
};
template <> struct Status<true> {
    using T = uint32_t;
    __device__ static T pack(uint32_t epoch, uint64_t flag, uint64_t v)
    {
        return ((epoch & 15u) << 28) | ((uint32_t)flag << 26) | (uint32_t)v;
    }
    __device__ static uint32_t flag(T s, uint32_t epoch) { return (s >> 28) == (epoch & 15u) ? (s >> 26) & 3u : 0u; }
    __device__ static uint64_t val(T s) { return s & ((1u << 26) - 1); }
};

// Samples of a batch sorted each on its own ("segmented"): every tile holds
// items of one sample; sample s's tiles are [t0[s], t0[s+1]), its items
// [pos[s], pos[s+1]) (tuples; windows for the generating pass).
struct SegTiles {
    const uint64_t *t0;
    const uint64_t *pos;
    int ns;  // 0: one segment, tiles of kTile items from 0
};

// sample of tile t: last s with t0[s] <= t
__device__ __forceinline__ int seg_of_tile(const SegTiles &S, uint64_t t)
{
    int lo = 0, hi = S.ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S.t0[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// items [a, e) of tile t (tiles of `tile` items, per sample when segmented)
__device__ __forceinline__ void tile_range(const SegTiles &S, uint64_t t, uint64_t tile, uint64_t n, uint64_t &a,
                                           uint64_t &e)
{
    if (S.ns == 0) {
        a = t * tile;
        e = min(a + tile, n);
        return;
    }
    const int s = seg_of_tile(S, t);
    a = S.pos[s] + (t - S.t0[s]) * tile;
    e = min(a + tile, S.pos[s + 1]);
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
    uint32_t mask;  // a key digit: (key >> shift) & mask
    const uint64_t *grp_first;  // n_grp + 1 first read ids of the samples
    int n_grp;
    const uint64_t *digit_base;
    void *status;
    uint32_t *tile_ctr;
    uint32_t epoch;
    const uint64_t *offs;  // counted passes: output index of each (tile, digit)'s first tuple, tile-major
    uint8_t *dig_out;      // if set: the next pass's digit of every tuple, at its output index
    int next_shift;
    uint32_t next_mask;
    int xcd_tiles;         // counted passes: consecutive tiles on one XCD (workgroups go round-robin to 8 XCDs)
    SegTiles seg;             // the input's tiles (scatter passes): per sample when segmented
    const uint32_t *chk_cnt;  // counted generating pass: the per-tile counts its offsets came from, checked
    uint32_t *chk_err;        // set when a tile's items disagree with them (the tile then writes nothing)
};

// Tile of workgroup b of a grid of g when the 8 XCDs (which take workgroups
// round-robin) each take one contiguous range of tiles, in order: the partial
// cache lines two neighbouring tiles write at each digit's run boundary then
// meet in one L2.
__device__ __forceinline__ uint64_t xcd_tile(uint64_t b, uint64_t g)
{
    const uint64_t x = b & 7, per = g >> 3, rem = g & 7;
    return x * per + (x < rem ? x : rem) + (b >> 3);
}

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

// digit counts of the passes in pass_mask (and of the sample digit when there are samples)
__global__ __launch_bounds__(kRadix) void sort_hist_kernel(const uint64_t *__restrict__ keys,
                                                           const uint8_t *__restrict__ adj,
                                                           const uint32_t *__restrict__ lab,
                                                           const uint64_t *__restrict__ grp_first, int n_grp,
                                                           uint64_t n, uint32_t pass_mask, KeyLayout kl,
                                                           unsigned long long *__restrict__ ghist)
{
    __shared__ uint32_t h[kPasses][kRadix];
    __shared__ uint64_t s_first[kRadix];  // the samples' first read ids (at most kRadix samples)
    for (int j = threadIdx.x; j < kPasses * kRadix; j += blockDim.x) (&h[0][0])[j] = 0;
    if (grp_first)
        for (int j = threadIdx.x; j < n_grp; j += blockDim.x) s_first[j] = grp_first[j];
    __syncthreads();
    const bool keyed = (pass_mask >> 1) & 0xFF;  // only the arrays a counted digit needs are read
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (pass_mask & 1) atomicAdd(&h[0][adj[i]], 1u);
        if (keyed) {
            const uint64_t k = keys[i];
#pragma unroll
            for (int b = 0; b < 8; b++)
                if ((pass_mask >> (1 + b)) & 1) atomicAdd(&h[1 + b][kl.digit(k, b)], 1u);
        }
        if (grp_first) atomicAdd(&h[9][group_of(s_first, n_grp, lab[i])], 1u);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kPasses * kRadix; j += blockDim.x) {
        const uint32_t v = (&h[0][0])[j];
        if (v) atomicAdd(&ghist[j], (unsigned long long)v);
    }
}

__device__ __forceinline__ uint32_t digit_of(uint64_t k, uint8_t a, int shift, uint32_t mask)
{
    return shift < 0 ? (uint32_t)a : (uint32_t)(k >> shift) & mask;
}

// The generating pass: its items are the reads' k-mer windows, tile t
// covering windows [t * tile, (t + 1) * tile) in read order; a window that
// holds an 'N' or lies outside the scan's buckets is not an item.  The
// tuples come out in the same order as an enumeration followed by this pass.
// What a tile of the generating pass needs before it stages its reads, in
// one 64-byte line so that it costs one round trip: the reads [ra, rb] that
// overlap its windows, the first window of read ra, the reads' bases
// [b0, b1) and their FASTQ sequence span [f0, f1).
struct TileInfo {
    uint64_t ra, rb, pa, b0, b1, f0, w0;  // w0: the tile's first window
    uint32_t flen, nw;                    // FASTQ span f0 + flen; windows of the tile
};
static_assert(sizeof(TileInfo) == 64, "one line");

struct GenPass {
    KParams G;
    const TileInfo *info;  // per tile
    uint64_t tiles;
    uint64_t npos;         // windows of all reads
};

constexpr int kGenReads = 256;  // reads of a tile staged in LDS (more: everything read from memory)

// read holding window q: last r with pos_off[r] <= q
__device__ __forceinline__ uint64_t read_of_window(const uint64_t *pos_off, uint64_t N, uint64_t q)
{
    uint64_t lo = 0, hi = N - 1;
    while (lo < hi) {
        const uint64_t mid = (lo + hi + 1) >> 1;
        if (pos_off[mid] <= q) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void tile_info_kernel(KParams G, uint64_t tile, uint64_t tiles, uint64_t npos, SegTiles S,
                                 TileInfo *__restrict__ info)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tiles) return;
    uint64_t w0, w1;
    tile_range(S, t, tile, npos, w0, w1);
    TileInfo ti;
    ti.ra = read_of_window(G.pos_off, G.N, w0);
    ti.rb = read_of_window(G.pos_off, G.N, w1 - 1);
    ti.pa = G.pos_off[ti.ra];
    ti.b0 = G.base_off[ti.ra];
    ti.b1 = G.base_off[ti.rb + 1];
    ti.f0 = G.seq_off[ti.ra];
    ti.flen = (uint32_t)(G.seq_off[ti.rb] + G.seq_len[ti.rb] - ti.f0);
    ti.w0 = w0;
    ti.nw = (uint32_t)(w1 - w0);
    info[t] = ti;
}

// How a pass finds each (tile, digit)'s global offset: decoupled look-back
// over 32- or 64-bit status words, or offsets counted beforehand (the
// digits were counted per tile from the digit bytes the previous pass wrote,
// and scanned; no tile waits on another).
enum { kLbCounted = 0, kLb32 = 1, kLb64 = 2 };

// waves per SIMD the tile shape's LDS allows: two 1024-lane tiles, or four 512 x 6 tiles, per CU
template <int kSB, int kSI, bool GEN, int LB>
__global__ __launch_bounds__(kSB, (kSB == 512 && kSI <= 6) ? 8 : 2 * kSB / 256) void onesweep_kernel(SortPass P,
                                                                                                      GenPass A)
{
    constexpr bool S32 = LB == kLb32;
    constexpr int kSW = kSB / kWave;
    constexpr int kTile = kSB * kSI;
    static_assert(!GEN || 2 * kGenReads * 8 + (3 * kGenReads + 2) * 4 <= kTile * 8, "read staging exceeds the tile");
    // one staging buffer (keys, then labels + adj) keeps LDS under half a CU,
    // so two tiles run per CU and one tile's memory phases overlap the other's
    // ranking and look-back
    __shared__ uint64_t s_buf[kTile];
    __shared__ uint8_t s_dig[kTile];
    __shared__ uint16_t s_cnt[kSW][kRadix];
    __shared__ uint32_t s_dstart[kRadix];  // digit starts in the tile
    __shared__ uint64_t s_gbase[kRadix];
    __shared__ uint32_t s_wsum[kRadix / kWave];
    __shared__ uint32_t s_tile, s_nvalid;

    using S = Status<S32>;
    typename S::T *const status = (typename S::T *)P.status;
    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    const unsigned lane = __lane_id();
    for (int j = tid; j < kSW * kRadix; j += kSB) (&s_cnt[0][0])[j] = 0;
    // the sample pass searches the samples' first read ids (at most kRadix) in
    // LDS: s_gbase is idle until the look-back
    const uint64_t *grp_first = P.grp_first;
    if (!GEN && P.shift == -2) {
        for (int j = tid; j < P.n_grp; j += kSB) s_gbase[j] = P.grp_first[j];
        grp_first = s_gbase;
    }
    if (LB != kLbCounted && tid == 0) s_tile = atomicAdd(P.tile_ctr, 1u);
    if (LB == kLbCounted && tid == 0)  // tiles do not wait on each other: any order
        s_tile = P.xcd_tiles ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    __syncthreads();
    const uint64_t tile = s_tile;
    uint64_t base = tile * kTile, lim = P.n;  // the tile's items [base, lim) of the input
    if (!GEN) tile_range(P.seg, tile, kTile, P.n, base, lim);

    uint64_t k[kSI];
    uint32_t l[kSI];
    uint32_t ap[(kSI + 3) / 4];  // adj bytes, four per register
    uint32_t dr[kSI];            // digit << 16 (kRadix: not an item), then | rank
#pragma unroll
    for (int i = 0; i < (kSI + 3) / 4; i++) ap[i] = 0;
    if constexpr (GEN) {
        // reads [ra, rb] overlap the tile's windows.  When they fit, their table,
        // their 2-bit and 'N' words and their FASTQ sequence bytes are staged in
        // the (idle) staging buffer, so the items read LDS only; else memory.
        const TileInfo ti = A.info[tile];
        const uint64_t ra = ti.ra, rb = ti.rb;
        const uint32_t nr = (uint32_t)(rb - ra + 1);
        const uint64_t pa = ti.pa;
        const uint64_t b0 = ti.b0, b1 = ti.b1;  // bases of the reads
        const uint64_t f0 = ti.f0, f1 = ti.f0 + ti.flen;  // their FASTQ span
        base = ti.w0;  // the tile's first window; its windows [base, base + nw)
        lim = ti.w0 + ti.nw;
        const uint64_t wp0 = b0 >> 5, np = b1 > b0 ? ((b1 - 1) >> 5) + 2 - wp0 : 0;
        const uint64_t wn0 = b0 >> 6, nn = b1 > b0 ? ((b1 - 1) >> 6) + 2 - wn0 : 0;
        const uint64_t fw0 = f0 >> 2, nf = (f1 + 3) / 4 - fw0;
        constexpr uint32_t kTab = 2 * 8 * kGenReads + 4 * (3 * kGenReads + 2);  // bytes of the read table
        const uint64_t mapo = (kTab + 8 * (np + nn) + 4 * nf + 7) & ~7ull;  // byte offset of the window map
        const bool staged = nr <= (uint32_t)kGenReads && mapo + kTile <= 8u * kTile;
        // staged, a read's first base and first sequence byte are 32-bit offsets
        // from the staged words (bases from wn0 * 64, bytes from fw0 * 4), so the
        // items index LDS with 32-bit arithmetic
        uint32_t *s_rbase = (uint32_t *)s_buf, *s_rso = (uint32_t *)(s_buf + kGenReads);
        uint32_t *s_rpos = (uint32_t *)(s_buf + 2 * kGenReads), *s_rlen = s_rpos + kGenReads + 1,
                 *s_rlab = s_rlen + kGenReads;
        uint64_t *s_pw = s_buf + kTab / 8, *s_nw = s_pw + np;
        uint32_t *s_fw = (uint32_t *)(s_nw + nn);
        if (staged) {
            // a lane's first entry of every array and its first kFQ FASTQ words
            // are loaded together (clamped indices, no branch), then stored; what
            // is left (long reads) follows in plain loops.  nr >= 1 (rb >= ra).
            static_assert(kGenReads < kSB, "one table entry per lane");
            constexpr int kFQ = 6;
            const uint32_t *fq32 = (const uint32_t *)A.G.fq;  // 4-aligned words inside the records
            const uint32_t jt = min((uint32_t)tid, nr), jr = min((uint32_t)tid, nr - 1);
            const uint64_t t_pos = A.G.pos_off[ra + jt];
            const uint64_t t_base = A.G.base_off[ra + jr], t_so = A.G.seq_off[ra + jr];
            const uint32_t t_len = A.G.seq_len[ra + jr];
            const uint64_t v_pw = np ? A.G.packed[wp0 + min((uint64_t)tid, np - 1)] : 0;
            const uint64_t v_nw = nn ? A.G.nmask[wn0 + min((uint64_t)tid, nn - 1)] : 0;
            uint32_t v_fw[kFQ];
#pragma unroll
            for (int q = 0; q < kFQ; q++) v_fw[q] = nf ? fq32[fw0 + min((uint64_t)(q * kSB + tid), nf - 1)] : 0;
            const uint32_t t_lab = label_of(A.G, ra + jr);  // last: it branches on the scan
            if ((uint32_t)tid <= nr) s_rpos[tid] = (uint32_t)(t_pos - pa);
            if ((uint32_t)tid < nr) {
                s_rbase[tid] = (uint32_t)(t_base - (wn0 << 6));
                s_rso[tid] = (uint32_t)(t_so - (fw0 << 2));
                s_rlen[tid] = t_len;
                s_rlab[tid] = t_lab;
            }
            if ((uint64_t)tid < np) s_pw[tid] = v_pw;
            if ((uint64_t)tid < nn) s_nw[tid] = v_nw;
#pragma unroll
            for (int q = 0; q < kFQ; q++)
                if ((uint64_t)(q * kSB + tid) < nf) s_fw[q * kSB + tid] = v_fw[q];
            for (uint32_t j = kSB + tid; j < np; j += kSB) s_pw[j] = A.G.packed[wp0 + j];
            for (uint32_t j = kSB + tid; j < nn; j += kSB) s_nw[j] = A.G.nmask[wn0 + j];
            for (uint32_t j = kFQ * kSB + tid; j < nf; j += kSB) s_fw[j] = fq32[fw0 + j];
        }
        __syncthreads();
        // window -> read of the tile (one byte each), so that the items need no search
        uint8_t *s_map = (uint8_t *)s_buf + mapo;
        const uint32_t q0 = (uint32_t)(base - pa);  // the tile's first window, relative to the first read
        if (staged) {
            // lane tid fills entries [tid * kSI, tid * kSI + kSI): the read of its first
            // window by binary search (the last read whose first window is at or before
            // it; reads without windows share the next read's), then onwards
            const uint32_t w0 = q0 + (uint32_t)tid * kSI;
            uint32_t lo = 0, hi = nr - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (s_rpos[mid] <= w0) lo = mid; else hi = mid - 1;
            }
#pragma unroll
            for (int i = 0; i < kSI; i++) {
                while (lo + 1 < nr && s_rpos[lo + 1] <= w0 + i) lo++;
                s_map[(uint32_t)tid * kSI + i] = (uint8_t)lo;
            }
            __syncthreads();
        }
        // the items, with the read data in LDS (ST) or in memory: two copies of
        // the loop, so that no access tests which
        auto items = [&](auto st_tag) {
            constexpr bool ST = decltype(st_tag)::value;
            auto rpos = [&](uint32_t j) -> uint32_t {
                if constexpr (ST) return s_rpos[j];
                else return (uint32_t)(A.G.pos_off[ra + j] - pa);
            };
            // staged: w and x relative (bases from wn0 * 64, 2-bit word wp0 is
            // word 2 * wn0 + dpw of that origin, bytes from fw0 * 4)
            const uint32_t dpw = (uint32_t)(wp0 - 2 * wn0);
            auto nword = [&](uint64_t w) -> uint64_t {
                if constexpr (ST) return s_nw[(uint32_t)w];
                else return A.G.nmask[w];
            };
            auto pword = [&](uint64_t w) -> uint64_t {
                if constexpr (ST) return s_pw[(uint32_t)w - dpw];
                else return A.G.packed[w];
            };
            auto fqb = [&](uint64_t x) -> uint8_t {
                if constexpr (ST) return ((const uint8_t *)s_fw)[(uint32_t)x];
                else return A.G.fq[x];
            };
            uint32_t lo = 0;  // read of the lane's last item: its items are 64 windows apart, so a few steps on
#pragma unroll
            for (int i = 0; i < kSI; i++) {
                const uint64_t idx = base + (uint64_t)wave * (kSI * kWave) + (uint64_t)i * kWave + lane;
                dr[i] = (uint32_t)kRadix << 16;  // not an item
                k[i] = 0;
                l[i] = 0;
                if (idx >= lim) continue;
                const uint32_t q = (uint32_t)(idx - pa);
                // the read holding window idx: last j with pos[j] <= q
                if constexpr (ST) {
                    lo = s_map[q - q0];
                } else {
                    for (int st = 0; st < 3 && lo + 1 < nr && rpos(lo + 1) <= q; st++) lo++;
                    if (lo + 1 < nr && rpos(lo + 1) <= q) {
                        uint32_t hi = nr - 1;
                        lo++;
                        while (lo < hi) {
                            const uint32_t mid = (lo + hi + 1) >> 1;
                            if (rpos(mid) <= q) lo = mid; else hi = mid - 1;
                        }
                    }
                }
                const uint64_t r = ra + lo;
                const uint32_t p = q - rpos(lo);
                uint64_t b, x;  // the window's first base and its first sequence byte (staged: 32-bit offsets)
                uint32_t L;
                if constexpr (ST) {
                    b = (uint32_t)(s_rbase[lo] + p);
                    x = (uint32_t)(s_rso[lo] + p);
                    L = s_rlen[lo];
                } else {
                    b = A.G.base_off[r] + p;
                    x = A.G.seq_off[r] + p;
                    L = A.G.seq_len[r];
                }
                uint64_t canon;
                bool fwd;
                if (!kmer_from(A.G.K, b, p, nword, pword, canon, fwd)) continue;
                const uint64_t bk = canon >> A.G.bshift;
                if (bk < A.G.lo || bk >= A.G.hi) continue;
                k[i] = canon;
                const uint8_t prev = p == 0 ? (uint8_t)'N' : fqb(x - 1);
                const uint8_t next = p == L - (uint32_t)A.G.K ? (uint8_t)'N' : fqb(x + A.G.K);
                const uint8_t a = adj_from(prev, next, fwd);
                ap[i / 4] |= (uint32_t)a << (8 * (i % 4));
                if constexpr (ST) l[i] = s_rlab[lo];
                else l[i] = label_of(A.G, r);
                dr[i] = digit_of(k[i], a, P.shift, P.mask) << 16;
            }
        };
        if (staged) {
            // Branch-free items from LDS, so that the compiler can overlap the
            // items' LDS reads: every item reads in-bounds entries (one past the
            // tile's windows takes the tile's first window and is dropped), the
            // window's K bits are the top of a 64-bit funnel over two words
            // (kmer_from's test and code, without its branches), and the
            // neighbour bytes are read at the window itself where 'N' stands in.
            const uint32_t dpw = (uint32_t)(wp0 - 2 * wn0), nwin = (uint32_t)(lim - base);
            const int K = A.G.K;
            const uint8_t *s_fb = (const uint8_t *)s_fw;
#pragma unroll
            for (int i = 0; i < kSI; i++) {
                const uint32_t iw = (uint32_t)(wave * (kSI * kWave) + i * kWave) + lane;
                const bool in = iw < nwin;
                const uint32_t lo = s_map[in ? iw : 0u];
                const uint32_t p = q0 + (in ? iw : 0u) - s_rpos[lo];
                const uint32_t b = s_rbase[lo] + p, x = s_rso[lo] + p, L = s_rlen[lo];
                const uint32_t o = b & 63, wn = b >> 6;
                const uint64_t nx = (s_nw[wn] << o) | ((s_nw[wn + 1] >> 1) >> (63 - o));
                const uint32_t o2 = 2 * (b & 31), wp = (b >> 5) - dpw;
                const uint64_t px = (s_pw[wp] << o2) | ((s_pw[wp + 1] >> 1) >> (63 - o2));
                uint64_t F = px >> (64 - 2 * K);
                const uint64_t R = revcomp(F, K);
                if (K == 32 && p > 0) F = 0;
                const bool fwd = F < R;
                const uint64_t canon = fwd ? F : R, bk = canon >> A.G.bshift;
                const bool valid = in && (nx >> (64 - K)) == 0 && bk >= A.G.lo && bk < A.G.hi;
                const bool first = p == 0, last = p == L - (uint32_t)K;
                const uint8_t pc = s_fb[first ? x : x - 1], nc = s_fb[last ? x : x + K];
                const uint8_t a = adj_from(first ? (uint8_t)'N' : pc, last ? (uint8_t)'N' : nc, fwd);
                const uint32_t lab = s_rlab[lo];
                k[i] = valid ? canon : 0;
                l[i] = valid ? lab : 0;
                ap[i / 4] |= (valid ? (uint32_t)a : 0u) << (8 * (i % 4));
                dr[i] = valid ? digit_of(canon, a, P.shift, P.mask) << 16 : (uint32_t)kRadix << 16;
            }
        } else {
            items(std::false_type{});
        }
    } else {
        // every item's loads first, so that a lane's loads are all in flight
        // together; then the digits (the sample pass's digit is a search loop,
        // which would otherwise hold each item's loads back until the previous
        // item's digit is known)
        uint32_t av[kSI];
#pragma unroll
        for (int i = 0; i < kSI; i++) {
            const uint64_t idx = base + (uint64_t)wave * (kSI * kWave) + (uint64_t)i * kWave + lane;
            k[i] = 0;
            l[i] = 0;
            av[i] = 0;
            if (idx < lim) {
                k[i] = P.keys_in[idx];
                l[i] = P.lab_in[idx];
                av[i] = P.adj_in[idx];
            }
        }
#pragma unroll
        for (int i = 0; i < kSI; i++) {
            const uint64_t idx = base + (uint64_t)wave * (kSI * kWave) + (uint64_t)i * kWave + lane;
            dr[i] = (uint32_t)kRadix << 16;  // invalid
            if (idx < lim) {
                ap[i / 4] |= av[i] << (8 * (i % 4));
                dr[i] = (P.shift == -2 ? group_of(grp_first, P.n_grp, l[i]) : digit_of(k[i], (uint8_t)av[i], P.shift, P.mask))
                        << 16;
            }
        }
    }
    // stable rank within the wave: item-major, lane-minor order
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int i = 0; i < kSI; i++) {
        const uint32_t d = dr[i] >> 16;
        const bool valid = d < (uint32_t)kRadix;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1;
            const uint64_t m = __ballot(valid && bit);
            peers &= bit ? m : ~m;
        }
        if (valid) {
            const uint32_t below = (uint32_t)__popcll(peers & lt);
            const uint32_t cnt = (uint32_t)__popcll(peers);
            const uint32_t old = s_cnt[wave][d];
            dr[i] = (d << 16) | (old + below);
            __builtin_amdgcn_wave_barrier();
            if (below == cnt - 1) s_cnt[wave][d] = (uint16_t)(old + cnt);
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
        if constexpr (LB != kLbCounted)
            st_agent(status + tile * kRadix + tid, S::pack(P.epoch, tile == 0 ? kStInc : kStAgg, total));
        uint32_t inc = total;
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t v = __shfl_up(inc, o, kWave);
            if (lane >= (unsigned)o) inc += v;
        }
        if (lane == kWave - 1) s_wsum[wave] = inc;
        s_dstart[tid] = inc - total;  // wave-local exclusive prefix; waves added below
    }
    __syncthreads();
    uint32_t ds = 0;
    if (tid < kRadix) {
        uint32_t woff = 0;
        for (int w = 0; w < wave; w++) woff += s_wsum[w];
        ds = s_dstart[tid] + woff;
        s_dstart[tid] = ds;
        if (tid == kRadix - 1) s_nvalid = ds + total;  // items of the tile
#pragma unroll
        for (int w = 0; w < kSW; w++) s_cnt[w][tid] = (uint16_t)(s_cnt[w][tid] + ds);  // tile position of wave w's first
    }
    __syncthreads();
    // tile-local position of every item, keys staged by it (the items' keys
    // and digits are dead afterwards, which leaves registers for the look-back)
#pragma unroll
    for (int i = 0; i < kSI; i++) {
        const uint32_t d = dr[i] >> 16;
        if (d < (uint32_t)kRadix) {
            const uint32_t rk = (dr[i] & 0xFFFF) + s_cnt[wave][d];
            s_buf[rk] = k[i];
            s_dig[rk] = (uint8_t)d;
            dr[i] = rk;  // the rank from here on
        } else {
            dr[i] = ~0u;
        }
    }
    if (LB == kLbCounted && tid < kRadix) {
        s_gbase[tid] = P.offs[tile * kRadix + tid] - ds;  // output index of tile position 0 of this digit
    }
    if (GEN && LB == kLbCounted) {
        // offsets counted from the histogram's windows: a tile whose items differ writes nothing
        const bool bad = tid < kRadix && P.chk_cnt[tile * kRadix + tid] != total;
        if (__syncthreads_or(bad)) {
            if (tid == 0) atomicOr(P.chk_err, 1u);
            return;
        }
    }
    if (LB != kLbCounted && tid < kRadix) {
        // decoupled look-back for digit tid
        uint64_t prefix = 0;
        if (tile != 0) {
            // kLookBack predecessors per step, their loads in flight together;
            // tile 0 always publishes an inclusive value, so the walk ends there at the latest
            int64_t j = (int64_t)tile - 1;
            bool done = false;
            while (!done) {
                typename S::T w[kLookBack];
#pragma unroll
                for (int q = 0; q < kLookBack; q++)
                    w[q] = j - q >= 0 ? ld_agent(status + (uint64_t)(j - q) * kRadix + tid) : 0;
#pragma unroll
                for (int q = 0; q < kLookBack; q++) {
                    if (done || j - q < 0) break;
                    typename S::T s = w[q];
                    uint32_t f = S::flag(s, P.epoch);
                    while (!f) {  // not published yet
                        s = ld_agent(status + (uint64_t)(j - q) * kRadix + tid);
                        f = S::flag(s, P.epoch);
                    }
                    prefix += S::val(s);
                    done = f == kStInc;
                }
                j -= kLookBack;
            }
            st_agent(status + tile * kRadix + tid, S::pack(P.epoch, kStInc, prefix + total));
        }
        s_gbase[tid] = P.digit_base[tid] + prefix - ds;  // output index of tile position 0 of this digit
    }
    __syncthreads();
    // tile position j = i * kSB + tid goes to output index gp[i]; the loops
    // have a fixed trip count and read LDS without a branch (j < kTile: the
    // reads stay in the arrays; past nvalid they are not used), so that a
    // lane's LDS reads are all issued before its stores (a loop over j waits
    // on two dependent LDS reads per step)
    const uint32_t nvalid = s_nvalid;
    uint64_t gp[kSI];
#pragma unroll
    for (int i = 0; i < kSI; i++) {
        const uint32_t j = i * kSB + tid;
        gp[i] = s_gbase[s_dig[j]] + j;
    }
    {
        uint64_t kv[kSI];  // every read before the first store: the stores might alias LDS for all the compiler knows
#pragma unroll
        for (int i = 0; i < kSI; i++) kv[i] = s_buf[i * kSB + tid];
#pragma unroll
        for (int i = 0; i < kSI; i++) {
            if (i * kSB + tid < (int)nvalid) {
                P.keys_out[gp[i]] = kv[i];
                if (P.dig_out) P.dig_out[gp[i]] = (uint8_t)((kv[i] >> P.next_shift) & P.next_mask);
            }
        }
    }
    __syncthreads();
    uint32_t *s_lab = (uint32_t *)s_buf;
    uint8_t *s_adj = (uint8_t *)(s_lab + kTile);
#pragma unroll
    for (int i = 0; i < kSI; i++) {
        if (dr[i] != ~0u) {
            s_lab[dr[i]] = l[i];
            s_adj[dr[i]] = (uint8_t)(ap[i / 4] >> (8 * (i % 4)));
        }
    }
    __syncthreads();
    uint32_t lv[kSI], av[kSI];
#pragma unroll
    for (int i = 0; i < kSI; i++) {
        lv[i] = s_lab[i * kSB + tid];
        av[i] = s_adj[i * kSB + tid];
    }
#pragma unroll
    for (int i = 0; i < kSI; i++) {
        if (i * kSB + tid < (int)nvalid) {
            P.lab_out[gp[i]] = lv[i];
            P.adj_out[gp[i]] = (uint8_t)av[i];
        }
    }
}

template <bool GEN, int LB>
static int launch_shape(mp_ctx *ctx, int SB, int SI, uint64_t tiles, const SortPass &P, const GenPass &A)
{
    if (SB == 512 && SI == 6)
        hipLaunchKernelGGL((onesweep_kernel<512, 6, GEN, LB>), dim3((unsigned)tiles), dim3(512), 0, ctx->stream, P, A);
    else if (SB == 512 && SI == 8)
        hipLaunchKernelGGL((onesweep_kernel<512, 8, GEN, LB>), dim3((unsigned)tiles), dim3(512), 0, ctx->stream, P, A);
    else if (SB == 512 && SI == 12)
        hipLaunchKernelGGL((onesweep_kernel<512, 12, GEN, LB>), dim3((unsigned)tiles), dim3(512), 0, ctx->stream, P,
                           A);
    else if (SB == 1024 && SI == 8)
        hipLaunchKernelGGL((onesweep_kernel<1024, 8, GEN, LB>), dim3((unsigned)tiles), dim3(1024), 0, ctx->stream, P,
                           A);
    else
        hipLaunchKernelGGL((onesweep_kernel<1024, 6, GEN, LB>), dim3((unsigned)tiles), dim3(1024), 0, ctx->stream, P,
                           A);
    MP_HIP(hipGetLastError());
    return 0;
}

template <bool GEN>
static int launch_onesweep(mp_ctx *ctx, int SB, int SI, uint64_t tiles, int lb, const SortPass &P, const GenPass &A)
{
    if (!grid_fits(tiles, SB))
        return ctx->fail(MP_EINVAL, "radix pass of %llu tiles exceeds one dispatch: use more scans",
                         (unsigned long long)tiles);
    if (lb == kLb32) return launch_shape<GEN, kLb32>(ctx, SB, SI, tiles, P, A);
    if (lb == kLb64) return launch_shape<GEN, kLb64>(ctx, SB, SI, tiles, P, A);
    return launch_shape<GEN, kLbCounted>(ctx, SB, SI, tiles, P, A);
}

// ---------------------------------------------------------------- counted passes
// A pass whose digits were written as one byte per tuple by the pass before
// (dig) finds its offsets without look-back: the digits are counted per tile
// (kCntG tiles per workgroup, also summed per such chunk), the chunk sums are
// scanned per digit from the digit's global base, and each tile's offsets are
// the chunk's plus the counts of the chunk's earlier tiles.  The extra bytes
// per tuple are one read (the count) and one write (the next pass's digit):
// no tile of the scatter kernel waits for another.
constexpr int kCntG = 16;

__global__ __launch_bounds__(256) void digit_count_kernel(const uint8_t *__restrict__ dig, uint64_t n, uint32_t tile,
                                                          uint64_t tiles, SegTiles S, uint32_t *__restrict__ cnt,
                                                          uint32_t *__restrict__ csum)
{
    __shared__ uint32_t h[kRadix];
    const uint64_t c = blockIdx.x, t0 = c * kCntG;
    uint32_t sum = 0;
    for (uint64_t t = t0; t < t0 + kCntG && t < tiles; t++) {
        h[threadIdx.x] = 0;
        __syncthreads();
        uint64_t a, e;
        tile_range(S, t, tile, n, a, e);
        // whole 8-byte words in the middle, single bytes at the ends
        const uint64_t am = min(e, (a + 7) & ~7ull), em = max(am, e & ~7ull);
        if (threadIdx.x < am - a) atomicAdd(&h[dig[a + threadIdx.x]], 1u);
        if (threadIdx.x < e - em) atomicAdd(&h[dig[em + threadIdx.x]], 1u);
        for (uint64_t i = am + threadIdx.x * 8; i < em; i += 256 * 8) {
            const uint64_t v = *(const uint64_t *)(dig + i);
#pragma unroll
            for (int b = 0; b < 8; b++) atomicAdd(&h[(v >> (8 * b)) & 0xFF], 1u);
        }
        __syncthreads();
        const uint32_t v = h[threadIdx.x];
        cnt[t * kRadix + threadIdx.x] = v;
        sum += v;
        __syncthreads();
    }
    csum[c * kRadix + threadIdx.x] = sum;
}

// the same digit counts from the keys themselves (byte `shift / 8` of each)
__global__ __launch_bounds__(256) void key_digit_count_kernel(const uint64_t *__restrict__ keys, uint64_t n, int shift,
                                                              uint32_t mask, uint32_t tile, uint64_t tiles, SegTiles S,
                                                              uint32_t *__restrict__ cnt, uint32_t *__restrict__ csum)
{
    __shared__ uint32_t h[kRadix];
    const uint64_t c = blockIdx.x, t0 = c * kCntG;
    uint32_t sum = 0;
    for (uint64_t t = t0; t < t0 + kCntG && t < tiles; t++) {
        h[threadIdx.x] = 0;
        __syncthreads();
        uint64_t a, e;
        tile_range(S, t, tile, n, a, e);
        for (uint64_t i = a + threadIdx.x; i < e; i += 256) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & mask], 1u);
        __syncthreads();
        const uint32_t v = h[threadIdx.x];
        cnt[t * kRadix + threadIdx.x] = v;
        sum += v;
        __syncthreads();
    }
    csum[c * kRadix + threadIdx.x] = sum;
}

// coff[c][d] = base[d] + sum of csum[c'][d] over c' < c; one workgroup per digit
__global__ __launch_bounds__(256) void chunk_scan_kernel(const uint32_t *__restrict__ csum, uint64_t nchunks,
                                                         const uint64_t *__restrict__ digit_base,
                                                         uint64_t *__restrict__ coff, uint64_t *__restrict__ tot)
{
    __shared__ uint64_t s_w[256 / kWave];
    const uint32_t d = blockIdx.x;
    const unsigned lane = __lane_id(), wave = threadIdx.x / kWave;
    uint64_t run = digit_base[d];
    for (uint64_t c0 = 0; c0 < nchunks; c0 += 256) {
        const uint64_t c = c0 + threadIdx.x;
        const uint64_t v = c < nchunks ? csum[c * kRadix + d] : 0;
        uint64_t inc = v;
        for (int o = 1; o < kWave; o <<= 1) {
            const uint64_t u = __shfl_up(inc, o, kWave);
            if (lane >= (unsigned)o) inc += u;
        }
        if (lane == kWave - 1) s_w[wave] = inc;
        __syncthreads();
        uint64_t woff = 0, tot_c = 0;
        for (int w = 0; w < 256 / kWave; w++) {
            woff += w < (int)wave ? s_w[w] : 0;
            tot_c += s_w[w];
        }
        if (c < nchunks) coff[c * kRadix + d] = run + woff + inc - v;
        run += tot_c;
        __syncthreads();
    }
    if (tot && threadIdx.x == 0) tot[d] = run - digit_base[d];  // tuples of digit d
}

// csum[c][d] = counts of digit d in the kCntG tiles of chunk c
__global__ __launch_bounds__(256) void chunk_sum_kernel(const uint32_t *__restrict__ cnt, uint64_t tiles,
                                                        uint32_t *__restrict__ csum)
{
    const uint64_t c = blockIdx.x, t0 = c * kCntG;
    uint32_t sum = 0;
    for (uint64_t t = t0; t < t0 + kCntG && t < tiles; t++) sum += cnt[t * kRadix + threadIdx.x];
    csum[c * kRadix + threadIdx.x] = sum;
}

// offs[t][d] = coff[chunk of t][d] + counts of d in the chunk's tiles before
// t, plus corr[sample of t][d] when the samples are sorted each on its own
__global__ __launch_bounds__(256) void tile_offsets_kernel(const uint32_t *__restrict__ cnt, uint64_t tiles,
                                                           const uint64_t *__restrict__ coff, SegTiles S,
                                                           const uint64_t *__restrict__ corr,
                                                           uint64_t *__restrict__ offs)
{
    const uint64_t c = blockIdx.x, t0 = c * kCntG;
    uint64_t run = coff[c * kRadix + threadIdx.x];
    for (uint64_t t = t0; t < t0 + kCntG && t < tiles; t++) {
        const uint64_t add = corr ? corr[(uint64_t)seg_of_tile(S, t) * kRadix + threadIdx.x] : 0;
        offs[t * kRadix + threadIdx.x] = run + add;
        run += cnt[t * kRadix + threadIdx.x];
    }
}

// Per-sample correction of segmented offsets.  With coff scanned from 0, the
// running count before tile t is P(t) (digit-major over all tiles); sample s
// holds P(t0[s+1]) - P(t0[s]) items of digit d, and its digit-d items go to
// pos[s] + (its items of smaller digits) + (P(t) - P(t0[s])).  One workgroup
// per sample, one lane per digit.
__device__ __forceinline__ uint64_t prefix_before(const uint32_t *cnt, const uint64_t *coff, const uint64_t *tot,
                                                  uint64_t tiles, uint64_t t, int d)
{
    if (t >= tiles) return tot[d];
    const uint64_t c0 = t / kCntG * kCntG;
    uint64_t v = coff[(t / kCntG) * kRadix + d];
    for (uint64_t u = c0; u < t; u++) v += cnt[u * kRadix + d];
    return v;
}

__global__ __launch_bounds__(256) void seg_fix_kernel(const uint32_t *__restrict__ cnt, const uint64_t *__restrict__ coff,
                                                      const uint64_t *__restrict__ tot, uint64_t tiles, SegTiles S,
                                                      uint64_t *__restrict__ corr)
{
    __shared__ uint64_t s_w[256 / kWave];
    const int sg = blockIdx.x, d = threadIdx.x;
    const unsigned lane = __lane_id(), wave = threadIdx.x / kWave;
    const uint64_t p0 = prefix_before(cnt, coff, tot, tiles, S.t0[sg], d);
    const uint64_t p1 = prefix_before(cnt, coff, tot, tiles, S.t0[sg + 1], d);
    const uint64_t v = p1 - p0;  // items of digit d in the sample
    uint64_t inc = v;
    for (int o = 1; o < kWave; o <<= 1) {
        const uint64_t u = __shfl_up(inc, o, kWave);
        if (lane >= (unsigned)o) inc += u;
    }
    if (lane == kWave - 1) s_w[wave] = inc;
    __syncthreads();
    uint64_t woff = 0;
    for (int w = 0; w < (int)wave; w++) woff += s_w[w];
    corr[(uint64_t)sg * kRadix + d] = S.pos[sg] + woff + inc - v - p0;
}

// tile shape: lanes x items.  Default from a sweep on MI355X at 2.4e9
// tuples: 1024 x 6 (two tiles per CU, 22.8 ms/pass) < 1024 x 8 (one tile per
// CU, 27.8) < 512 x 8 (32.8) < 512 x 16 (34.3).
static void sort_shape(const mp_ctx *ctx, int *SB, int *SI)
{
    int shape = ctx->cfg.sort_tile;
    if (const char *e = getenv("MP_SORT_TILE")) shape = atoi(e);
    switch (shape) {
    case 3072: *SB = 512; *SI = 6; break;
    case 4096: *SB = 512; *SI = 8; break;
    case 6144: *SB = 512; *SI = 12; break;
    case 8192: *SB = 1024; *SI = 8; break;
    default: *SB = 1024; *SI = 6; break;  // 6144 tuples on 1024 lanes
    }
}

uint64_t sort_tile_items(const mp_ctx *ctx)
{
    int SB = 1024, SI = 6;
    sort_shape(ctx, &SB, &SI);
    return (uint64_t)SB * SI;
}

static uint64_t kl_id(const KeyLayout &L)
{
    uint64_t v = 0;
    for (int b = 0; b < 8; b++) v = v * 131 + L.shift[b] * 16 + L.bits[b];
    return v;
}

int radix_sort_generic(mp_ctx *ctx, DevArray *keys, DevArray *lab, DevArray *adj, int *cur, uint64_t n,
                       uint32_t pass_mask, int *passes_out, const uint64_t *d_grp_first, int n_grp,
                       uint64_t *grp_base, const GenSource *gen, const SortSegments *segs, const KeyLayout *kl)
{
    const KeyLayout KL = kl ? *kl : byte_layout();
    if (segs && segs->ns > 0 && (!gen || !gen->gcnt || !gen->gcnt_seg))
        return ctx->fail(MP_EINVAL, "a segmented generated sort needs the per-sample tile counts");
    *passes_out = 0;
    if (gen ? gen->npos == 0 : n <= 1) return 0;
    // histograms of every digit
    MP_TRY(ensure(ctx, ctx->hist, sizeof(uint64_t) * (kPasses * kRadix + 64)));
    unsigned long long *gh = (unsigned long long *)ctx->hist.p;
    uint32_t *tile_ctr = (uint32_t *)(gh + kPasses * kRadix);
    MP_TRY(fill_u32(ctx, gh, 0, 2 * kPasses * kRadix + 2 * 32));
    hipEventRecord(ctx->pev[0], ctx->stream);
    if (!gen) {
        uint64_t blocks = (n + kRadix * 16 - 1) / (kRadix * 16);
        if (blocks > 2048) blocks = 2048;
        hipLaunchKernelGGL(sort_hist_kernel, dim3((unsigned)blocks), dim3(kRadix), 0, ctx->stream,
                           (const uint64_t *)keys[*cur].p, (const uint8_t *)adj[*cur].p,
                           (const uint32_t *)lab[*cur].p, d_grp_first, n_grp, n, pass_mask, KL, gh);
        MP_HIP(hipGetLastError());
    }
    hipEventRecord(ctx->pev[1], ctx->stream);
    uint64_t h[kPasses * kRadix];
    if (gen) {
        memcpy(h, gen->h, sizeof(h));
    } else {
        MP_HIP(hipMemcpyAsync(h, gh, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        MP_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (!d_grp_first) h[9 * kRadix] = n;  // no samples: one constant digit
    // digit base offsets of the passes that do work
    uint64_t base[kPasses][kRadix], dmax[kPasses];
    bool run[kPasses];
    for (int p = 0; p < kPasses; p++) {
        uint64_t acc = 0, mx = 0;
        for (int r = 0; r < kRadix; r++) {
            base[p][r] = acc;
            acc += h[p * kRadix + r];
            if (h[p * kRadix + r] > mx) mx = h[p * kRadix + r];
        }
        if (((pass_mask >> p) & 1) && acc != n)
            return ctx->fail(MP_EDEVICE, "sort histogram total %llu != %llu", (unsigned long long)acc,
                             (unsigned long long)n);
        run[p] = ((pass_mask >> p) & 1) && mx != n;
        dmax[p] = mx;
    }
    // the generating pass: the first pass that does work, else the lowest of the mask
    int gen_p = -1;
    if (gen) {
        for (int p = 0; p < kPasses && gen_p < 0; p++)
            if (run[p]) gen_p = p;
        if (gen_p < 0) gen_p = __builtin_ctz(pass_mask);
        if (gen_p < 1 || gen_p > 8) return ctx->fail(MP_EINVAL, "generated sorts take k-mer byte passes only");
        run[gen_p] = true;
    }
    if (n_grp > 0 && grp_base) {  // tuple offset of every sample after the sort
        for (int g = 0; g < n_grp; g++) grp_base[g] = base[9][g];
        grp_base[n_grp] = n;
    }
    int SB = 1024, SI = 6;
    sort_shape(ctx, &SB, &SI);
    const uint64_t kTile = (uint64_t)SB * SI;
    uint64_t tiles = (n + kTile - 1) / kTile;
    uint64_t gtiles = gen ? (gen->npos + kTile - 1) / kTile : 0;
    // segmented: the tiles of every sample (tuples; windows of the generating pass)
    const int ns = segs ? segs->ns : 0;
    std::vector<uint64_t> seg_host;  // tt0[ns + 1], tpos[ns + 1], gt0[ns + 1], wpos[ns + 1]
    SegTiles segT{}, segG{}, segGout{};
    uint64_t *d_seg = nullptr;
    if (ns > 0) {
        seg_host.assign(4 * (size_t)(ns + 1), 0);
        uint64_t *tt0 = seg_host.data(), *tp = tt0 + (ns + 1), *gt0 = tp + (ns + 1), *wp = gt0 + (ns + 1);
        for (int i = 0; i <= ns; i++) {
            tp[i] = segs->tpos[i];
            wp[i] = gen ? segs->wpos[i] : 0;
        }
        for (int i = 0; i < ns; i++) {
            tt0[i + 1] = tt0[i] + (tp[i + 1] - tp[i] + kTile - 1) / kTile;
            gt0[i + 1] = gt0[i] + (wp[i + 1] - wp[i] + kTile - 1) / kTile;
        }
        if (tp[ns] != n) return ctx->fail(MP_EINVAL, "segments hold %llu tuples, the sort %llu",
                                          (unsigned long long)tp[ns], (unsigned long long)n);
        tiles = tt0[ns];
        if (gen) gtiles = gt0[ns];
        if (gen && gtiles != gen->gcnt_tiles)
            return ctx->fail(MP_EINVAL, "per-sample tiles differ from the histogram's");
    }
    const uint64_t st_bytes = (tiles > gtiles ? tiles : gtiles) * kRadix * sizeof(uint64_t);
    if (ctx->status.bytes < st_bytes) {
        MP_TRY(ensure(ctx, ctx->status, st_bytes));
        MP_TRY(fill_u32(ctx, ctx->status.p, 0, ctx->status.bytes / 4));
        ctx->epoch = 0;
    }
    // digit bases, segment tables and the generating pass's tile table in one
    // buffer that stays with the ctx (a captured pass chain reads them there)
    const uint64_t aux_seg = sizeof(base), aux_ti = (aux_seg + seg_host.size() * 8 + 63) & ~63ull;
    MP_TRY(ensure(ctx, ctx->sortaux, aux_ti + (gen ? gtiles * sizeof(TileInfo) : 0) + 64));
    uint64_t *d_base = (uint64_t *)ctx->sortaux.p;
    TileInfo *d_tile_info = gen ? (TileInfo *)((uint8_t *)ctx->sortaux.p + aux_ti) : nullptr;
    MP_HIP(hipMemcpyAsync(d_base, base, sizeof(base), hipMemcpyHostToDevice, ctx->stream));
    if (ns > 0) {
        d_seg = (uint64_t *)((uint8_t *)ctx->sortaux.p + aux_seg);
        MP_HIP(hipMemcpyAsync(d_seg, seg_host.data(), seg_host.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                              ctx->stream));
        segT = SegTiles{d_seg, d_seg + (ns + 1), ns};
        segG = SegTiles{d_seg + 2 * (ns + 1), d_seg + 3 * (ns + 1), ns};
        segGout = SegTiles{d_seg + 2 * (ns + 1), d_seg + (ns + 1), ns};
    }
    GenPass A{};
    if (gen) {
        hipLaunchKernelGGL(tile_info_kernel, dim3((unsigned)((gtiles + 255) / 256)), dim3(256), 0, ctx->stream,
                           gen->P, kTile, gtiles, gen->npos, segG, d_tile_info);
        MP_HIP(hipGetLastError());
        A.G = gen->P;
        A.info = d_tile_info;
        A.tiles = gtiles;
        A.npos = gen->npos;
    }
    int executed = 0;
    // The passes after the first take their offsets from digit counts, their
    // tiles in one contiguous range per XCD.  Measured at 2.4e9 tuples (k = 31,
    // 8-bit digits, 6144-tuple tiles), per scatter pass: look-back 21.2 ms;
    // counted, tiles in launch order 27.3 (+ 1.2 counting); counted, tiles by
    // XCD 17.4 (+ 1.2); counted from the keys instead of digit bytes, tiles by
    // XCD 16.4 (+ 4.0).  The partial cache lines at the run boundaries of two
    // neighbouring tiles are the cost: in one L2 they merge.
    // MP_SORT_LOOKBACK=1 takes every pass by look-back, MP_SORT_XCD=0 leaves the
    // tiles in launch order, MP_SORT_STATUS64 forces 64-bit status words on every
    // look-back pass (tests run the small inputs through the format the
    // largest inputs need).
    const bool force64 = getenv("MP_SORT_STATUS64") != nullptr;
    const char *lbe = getenv("MP_SORT_LOOKBACK"), *xe = getenv("MP_SORT_XCD");
    const bool counted = !(lbe && atoi(lbe) == 1);
    const bool count_keys = false;  // first passes of enumerated sorts count from the keys (below)
    const bool xcd = !(xe && atoi(xe) == 0);
    std::vector<int> plist;
    for (int p = 0; p < kPasses; p++)
        if (run[p]) plist.push_back(p);
    // the generating pass takes counted offsets too when the m-mer histogram
    // counted its digits per tile (a scan over every k-mer, same tile size)
    const bool gen_counted = counted && gen && gen->gcnt && gen_p == gen->gcnt_pass && gen->gcnt_tiles == gtiles &&
                             (ns > 0) == gen->gcnt_seg;
    if (ns > 0 && !gen_counted) return ctx->fail(MP_EINVAL, "a segmented sort takes counted offsets only");
    // counted passes: the digit bytes (n, padded to whole words) and the count/offset tables
    const uint64_t ctiles = tiles > gtiles ? tiles : gtiles;
    const uint64_t nchunks = (tiles + kCntG - 1) / kCntG, cchunks = (ctiles + kCntG - 1) / kCntG;
    uint8_t *d_dig = nullptr;
    uint32_t *d_cnt = nullptr, *d_csum = nullptr;
    uint64_t *d_coff = nullptr, *d_offs = nullptr, *d_tot = nullptr, *d_zero = nullptr, *d_corr = nullptr;
    if (counted && (plist.size() > 1 || gen_counted)) {
        MP_TRY(ensure(ctx, ctx->dig, (n + 15) & ~7ull));
        const uint64_t o_csum = ctiles * kRadix * 4, o_coff = (o_csum + cchunks * kRadix * 4 + 7) & ~7ull,
                       o_offs = o_coff + cchunks * kRadix * 8, o_tot = o_offs + ctiles * kRadix * 8,
                       o_zero = o_tot + kRadix * 8, o_corr = o_zero + kRadix * 8;
        MP_TRY(ensure(ctx, ctx->rts, o_corr + (uint64_t)ns * kRadix * 8 + 64));
        d_dig = (uint8_t *)ctx->dig.p;
        d_cnt = (uint32_t *)ctx->rts.p;
        d_csum = (uint32_t *)((uint8_t *)ctx->rts.p + o_csum);
        d_coff = (uint64_t *)((uint8_t *)ctx->rts.p + o_coff);
        d_offs = (uint64_t *)((uint8_t *)ctx->rts.p + o_offs);
        d_tot = (uint64_t *)((uint8_t *)ctx->rts.p + o_tot);
        d_zero = (uint64_t *)((uint8_t *)ctx->rts.p + o_zero);
        d_corr = (uint64_t *)((uint8_t *)ctx->rts.p + o_corr);
        if (ns > 0) MP_TRY(fill_u32(ctx, d_zero, 0, 2 * kRadix));
    }
    // the generating pass's offsets, checked against the digit totals before any tuple is written
    bool gen_offs = false;
    if (gen_counted && d_offs) {
        const uint64_t gchunks = (gtiles + kCntG - 1) / kCntG;
        hipEventRecord(ctx->cev[kPasses], ctx->stream);
        hipLaunchKernelGGL(chunk_sum_kernel, dim3((unsigned)gchunks), dim3(256), 0, ctx->stream, gen->gcnt, gtiles,
                           d_csum);
        // segmented: counts scanned from 0, then every sample's items from its own position
        hipLaunchKernelGGL(chunk_scan_kernel, dim3(kRadix), dim3(256), 0, ctx->stream, (const uint32_t *)d_csum,
                           gchunks, (const uint64_t *)(ns > 0 ? d_zero : d_base + gen_p * kRadix), d_coff, d_tot);
        if (ns > 0)
            hipLaunchKernelGGL(seg_fix_kernel, dim3(ns), dim3(256), 0, ctx->stream, gen->gcnt,
                               (const uint64_t *)d_coff, (const uint64_t *)d_tot, gtiles, segGout, d_corr);
        hipLaunchKernelGGL(tile_offsets_kernel, dim3((unsigned)gchunks), dim3(256), 0, ctx->stream, gen->gcnt, gtiles,
                           (const uint64_t *)d_coff, segGout, (const uint64_t *)(ns > 0 ? d_corr : nullptr), d_offs);
        MP_HIP(hipGetLastError());
        uint64_t tot[kRadix];
        MP_HIP(hipMemcpyAsync(tot, d_tot, sizeof(tot), hipMemcpyDeviceToHost, ctx->stream));
        MP_HIP(hipStreamSynchronize(ctx->stream));
        gen_offs = true;
        for (int d = 0; d < kRadix; d++)
            if (tot[d] != h[gen_p * kRadix + d]) gen_offs = false;  // not the scan's tuples: look-back
    }
    bool counted_pass[kPasses] = {};
    const int cur0 = *cur;
    // The pass chain.  launch = false only advances the host state (the
    // buffers and passes) for a replayed graph; timing records events per pass.
    auto chain = [&](bool launch, bool timing) -> int {
    bool dig_valid = false;  // d_dig holds the digits of the pass about to run
    *cur = cur0;
    executed = 0;
    for (size_t pi = 0; pi < plist.size(); pi++) {
        const int p = plist[pi];
        const int q = pi + 1 < plist.size() ? plist[pi + 1] : -1;
        // a pass with no digit bytes before it (the first of an enumerated sort) counts from the keys
        const bool from_keys = d_dig && !dig_valid && pi == 0 && p >= 1 && p <= 8 && gen_p < 0;
        const bool this_counted =
            (d_dig && (dig_valid || from_keys || count_keys) && p != gen_p && p >= 1 && p <= 8) ||
            (p == gen_p && gen_offs);
        const bool next_counted = d_dig && q >= 1 && q <= 8;
        // status words: 32-bit when every digit's count fits 26 bits; the array is
        // cleared when the format changes or the epoch wraps (4 bits for 32-bit words)
        const bool s32 = dmax[p] < (1ull << 26) && !force64;
        const int lb = this_counted ? kLbCounted : s32 ? kLb32 : kLb64;
        if (!this_counted && launch) {
            ++ctx->epoch;
            if (ctx->epoch == 0 || ctx->status_fmt != (s32 ? 1 : 2) || (s32 && (ctx->epoch & 15) == 0)) {
                MP_TRY(fill_u32(ctx, ctx->status.p, 0, ctx->status.bytes / 4));
                if (ctx->epoch == 0) ctx->epoch = 1;
                ctx->status_fmt = s32 ? 1 : 2;
            }
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
        P.shift = p == 0 ? -1 : p == 9 ? -2 : (int)KL.shift[p - 1];
        P.mask = p >= 1 && p <= 8 ? (1u << KL.bits[p - 1]) - 1u : 0xFFu;
        P.grp_first = d_grp_first;
        P.n_grp = n_grp;
        P.digit_base = d_base + p * kRadix;
        P.status = ctx->status.p;
        P.tile_ctr = tile_ctr + executed;
        P.epoch = ctx->epoch;
        P.offs = d_offs;
        P.dig_out = next_counted ? d_dig : nullptr;
        P.next_shift = next_counted ? (int)KL.shift[q - 1] : 0;
        P.next_mask = next_counted ? (1u << KL.bits[q - 1]) - 1u : 0xFFu;
        P.xcd_tiles = xcd ? 1 : 0;
        P.seg = segT;
        P.chk_cnt = p == gen_p && gen_offs ? gen->gcnt : nullptr;
        P.chk_err = (uint32_t *)(tile_ctr + kPasses + 1);
        if (this_counted && p != gen_p && launch) {
            if (timing) hipEventRecord(ctx->cev[executed], ctx->stream);
            if (count_keys || from_keys)
                hipLaunchKernelGGL(key_digit_count_kernel, dim3((unsigned)nchunks), dim3(256), 0, ctx->stream,
                                   (const uint64_t *)keys[src].p, n, (int)KL.shift[p - 1],
                                   (1u << KL.bits[p - 1]) - 1u, (uint32_t)kTile, tiles, segT, d_cnt, d_csum);
            else
                hipLaunchKernelGGL(digit_count_kernel, dim3((unsigned)nchunks), dim3(256), 0, ctx->stream,
                                   (const uint8_t *)d_dig, n, (uint32_t)kTile, tiles, segT, d_cnt, d_csum);
            MP_HIP(hipGetLastError());
            hipLaunchKernelGGL(chunk_scan_kernel, dim3(kRadix), dim3(256), 0, ctx->stream, (const uint32_t *)d_csum,
                               nchunks, (const uint64_t *)(ns > 0 ? d_zero : P.digit_base), d_coff, d_tot);
            MP_HIP(hipGetLastError());
            if (ns > 0)
                hipLaunchKernelGGL(seg_fix_kernel, dim3(ns), dim3(256), 0, ctx->stream, (const uint32_t *)d_cnt,
                                   (const uint64_t *)d_coff, (const uint64_t *)d_tot, tiles, segT, d_corr);
            hipLaunchKernelGGL(tile_offsets_kernel, dim3((unsigned)nchunks), dim3(256), 0, ctx->stream,
                               (const uint32_t *)d_cnt, tiles, (const uint64_t *)d_coff, segT,
                               (const uint64_t *)(ns > 0 ? d_corr : nullptr), d_offs);
            MP_HIP(hipGetLastError());
        }
        counted_pass[executed] = this_counted && p != gen_p;
        if (launch) {
            if (timing) hipEventRecord(ctx->pev[2 + 2 * executed], ctx->stream);
            if (p == gen_p)
                MP_TRY(launch_onesweep<true>(ctx, SB, SI, gtiles, lb, P, A));
            else
                MP_TRY(launch_onesweep<false>(ctx, SB, SI, tiles, lb, P, A));
            if (timing) hipEventRecord(ctx->pev[3 + 2 * executed], ctx->stream);
        }
        dig_valid = next_counted;
        *cur = dst;
        executed++;
    }
    return 0;
    };
    // A chain of counted passes only (no look-back status epoch in its
    // arguments) may be captured in a HIP graph and replayed while the sort's
    // shape and buffers stay the same (mp_cfg.sort_graph).
    bool graphable = ctx->cfg.sort_graph != 0;
    {
        bool dv = false;
        for (size_t pi = 0; pi < plist.size() && graphable; pi++) {
            const int p = plist[pi];
            const int q = pi + 1 < plist.size() ? plist[pi + 1] : -1;
            const bool from_keys = d_dig && !dv && pi == 0 && p >= 1 && p <= 8 && gen_p < 0;
            const bool tc = (d_dig && (dv || from_keys) && p != gen_p && p >= 1 && p <= 8) || (p == gen_p && gen_offs);
            graphable = tc;
            dv = d_dig && q >= 1 && q <= 8;
        }
    }
    if (graphable) {
        const uint64_t key[] = {n, tiles, gtiles, (uint64_t)SB, (uint64_t)SI, (uint64_t)ns, (uint64_t)gen_p,
                                (uint64_t)pass_mask, (uint64_t)cur0, (uint64_t)xcd, kl_id(KL), (uint64_t)(uintptr_t)keys[0].p,
                                (uint64_t)(uintptr_t)keys[1].p, (uint64_t)(uintptr_t)lab[0].p,
                                (uint64_t)(uintptr_t)lab[1].p, (uint64_t)(uintptr_t)adj[0].p,
                                (uint64_t)(uintptr_t)adj[1].p, (uint64_t)(uintptr_t)d_dig, (uint64_t)(uintptr_t)d_cnt,
                                (uint64_t)(uintptr_t)d_offs, (uint64_t)(uintptr_t)d_base,
                                (uint64_t)(uintptr_t)d_tile_info, (uint64_t)(uintptr_t)d_seg,
                                (uint64_t)(uintptr_t)tile_ctr, (uint64_t)(uintptr_t)(gen ? gen->gcnt : nullptr),
                                gen ? gen->P.lo : 0, gen ? gen->P.hi : 0, gen ? (uint64_t)gen->P.scan0 : 0,
                                (uint64_t)(uintptr_t)(gen ? gen->P.labels : nullptr), gen ? gen->P.N : 0,
                                (uint64_t)(uintptr_t)(gen ? gen->P.packed : nullptr),
                                (uint64_t)(uintptr_t)(gen ? gen->P.fq : nullptr), gen ? gen->npos : 0};
        const std::vector<uint64_t> kv(key, key + sizeof(key) / sizeof(key[0]));
        hipEventRecord(ctx->pev[2], ctx->stream);
        constexpr size_t kGraphs = 4;  // chains kept; the oldest goes first
        hipGraphExec_t exec = nullptr;
        for (size_t i = 0; i < ctx->sort_graphs.size(); i++)
            if (ctx->sort_graphs[i].first == kv) {
                exec = ctx->sort_graphs[i].second;
                std::rotate(ctx->sort_graphs.begin() + i, ctx->sort_graphs.begin() + i + 1, ctx->sort_graphs.end());
                break;
            }
        if (exec) {
            MP_TRY(chain(false, false));
            ctx->sort_graph_replays++;
            ctx->sort_graph_last = 2;
        } else {
            if (ctx->sort_graphs.size() >= kGraphs) {
                (void)hipGraphExecDestroy(ctx->sort_graphs.front().second);
                ctx->sort_graphs.erase(ctx->sort_graphs.begin());
            }
            hipGraph_t g = nullptr;
            MP_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
            const int rc = chain(true, false);
            const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (ec != hipSuccess) return ctx->fail(MP_EDEVICE, "capture of the radix passes: %s", hipGetErrorString(ec));
            const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) return ctx->fail(MP_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ei));
            ctx->sort_graphs.emplace_back(kv, exec);
            ctx->sort_graph_captures++;
            ctx->sort_graph_last = 1;
        }
        MP_HIP(hipGraphLaunch(exec, ctx->stream));
        hipEventRecord(ctx->pev[3], ctx->stream);
    } else {
        MP_TRY(chain(true, true));
        ctx->sort_graph_last = 0;
    }

    uint32_t chk_err = 0;
    if (gen_offs) MP_HIP(hipMemcpyAsync(&chk_err, tile_ctr + kPasses + 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    if (chk_err)
        return ctx->fail(MP_EDEVICE, "generating pass: a tile's k-mers differ from the histogram's counts for it");
    float tot = 0, ms = 0;
    ctx->sort_gen_ms = 0;
    ctx->sort_count_ms = 0;
    if (graphable) {  // one graph: its time spread over its passes
        hipEventElapsedTime(&ms, ctx->pev[2], ctx->pev[3]);
        ctx->sort_pass_avg = executed ? ms / executed : 0.f;
        hipEventElapsedTime(&ctx->sort_hist_ms, ctx->pev[0], ctx->pev[1]);
        *passes_out = executed;
        return 0;
    }
    if (gen_counted && d_offs) {  // the generating pass's offsets (from the histogram's counts)
        hipEventElapsedTime(&ms, ctx->cev[kPasses], ctx->pev[2]);
        ctx->sort_count_ms += ms;
    }
    for (int e = 0; e < executed; e++) {
        if (counted_pass[e]) {
            hipEventElapsedTime(&ms, ctx->cev[e], ctx->pev[2 + 2 * e]);
            ctx->sort_count_ms += ms;
        }
        hipEventElapsedTime(&ms, ctx->pev[2 + 2 * e], ctx->pev[3 + 2 * e]);
        if (gen && e == 0) ctx->sort_gen_ms = ms;  // the generating pass, timed apart
        else tot += ms;
    }
    const int plain = executed - (gen && executed ? 1 : 0);
    ctx->sort_pass_avg = plain ? tot / plain : 0.f;
    hipEventElapsedTime(&ctx->sort_hist_ms, ctx->pev[0], ctx->pev[1]);
    *passes_out = executed;
    return 0;
}

// ---------------------------------------------------------------- hybrid order
// Global passes on the k-mer's high bytes only, then one local pass: after
// the high passes (stable, from generation order) tuples with equal high bits
// -- a "group" -- are contiguous and in generation order, and with about 30
// high bits a group holds a few tuples.  The local pass gives each tuple its
// rank inside its group by (k-mer, position) and writes it there: the same
// array the low-byte passes would have produced, for the cost of one pass.
// A workgroup owns the groups that start in its 1024-tuple window: the window
// and the next 256 positions are loaded into LDS at once, group starts are an
// LDS bitmap, and the window's last group may take up to 1024 positions past
// the window; a longer group goes to a list and is sorted by its low bytes
// with the global passes on its own range.
#ifndef MP_LOCAL_WIN
#define MP_LOCAL_WIN 1024
#endif
constexpr int kLWin = MP_LOCAL_WIN, kLMax = MP_LOCAL_WIN, kLCap = kLWin + kLMax, kLThreads = 256;
constexpr int kLMarg = 256;  // positions loaded past the window for the next group start

struct LocalPass {  // the tuple arrays, sorted in place
    uint64_t *keys;
    uint32_t *lab;
    uint8_t *adj;
    uint64_t n;
    int shift;                  // high bits: key >> shift
    uint64_t *ovf;              // starts of groups longer than kLMax
    unsigned long long *novf;
    uint32_t *moved;            // per wave: tuples written elsewhere than they were (summed afterwards)
    const uint64_t *seg;        // samples sorted each on its own: their first tuples (nseg + 1), group starts too
    int nseg;
};

// first sample start >= i (seg[nseg] = n)
__device__ __forceinline__ uint64_t next_seg_start(const uint64_t *seg, int nseg, uint64_t i)
{
    int lo = 0, hi = nseg;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (seg[mid] < i) lo = mid + 1; else hi = mid;
    }
    return seg[lo];
}

__device__ __forceinline__ bool group_start(const LocalPass &L, uint64_t i)
{
    if (i == 0 || (L.keys[i] >> L.shift) != (L.keys[i - 1] >> L.shift)) return true;
    return L.nseg && next_seg_start(L.seg, L.nseg, i) == i;
}

__global__ __launch_bounds__(kLThreads, 8) void local_rank_kernel(LocalPass L)
{
    // keys of [w0, w0 + kLCap) relative to w0, then the order keys of the groups
    __shared__ uint64_t s_key[kLCap + 4];  // 4 past the end: the rank loop reads whole steps of 4
    __shared__ uint64_t s_st[kLCap / kWave];  // group starts, bit j % 64 of word j / 64 for w0 + j
    __shared__ unsigned long long s_r1;
    __shared__ uint64_t s_prev;
    const uint64_t w0 = (uint64_t)blockIdx.x * kLWin, w1 = min(w0 + kLWin, L.n);
    const uint64_t lim = min(w1 + kLMax, L.n);
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    // the window and a margin after it, in one coalesced load; group starts
    // from LDS (the first position of the array is a start)
    const uint32_t nld = (uint32_t)(min(w1 + kLMarg, lim) - w0);
    {
        // a fixed trip count and no branch (positions past nld load the last
        // one, nld >= 1), so that a lane's loads are all issued before its stores
        constexpr int kLd = (kLWin + kLMarg) / kLThreads;
        static_assert(kLd * kLThreads == kLWin + kLMarg, "window load");
        uint64_t v[kLd];
#pragma unroll
        for (int q = 0; q < kLd; q++) v[q] = L.keys[w0 + min((uint32_t)(q * kLThreads + threadIdx.x), nld - 1)];
#pragma unroll
        for (int q = 0; q < kLd; q++) {
            const uint32_t j = q * kLThreads + threadIdx.x;
            if (j < nld) s_key[j] = v[q];
        }
    }
    if (threadIdx.x == 0) {
        s_prev = w0 ? L.keys[w0 - 1] : 0;
        s_r1 = lim;
    }
    __syncthreads();
    for (uint32_t j0 = wave * kWave; j0 < nld; j0 += kLThreads) {
        const uint32_t j = j0 + lane;
        const bool st = j < nld && (w0 + j == 0 || (s_key[j] >> L.shift) != ((j ? s_key[j - 1] : s_prev) >> L.shift));
        const uint64_t m = __ballot(st);
        if (lane == 0) s_st[j0 / kWave] = m;
    }
    __syncthreads();
    if (L.nseg && threadIdx.x == 0) {  // the samples' first tuples start groups too
        for (uint64_t p = next_seg_start(L.seg, L.nseg, w0); p < w0 + nld; p = next_seg_start(L.seg, L.nseg, p + 1))
            s_st[(p - w0) / kWave] |= 1ull << ((p - w0) % kWave);
    }
    __syncthreads();
    const uint32_t e0 = (uint32_t)(w1 - w0), nwl = (nld + kWave - 1) / kWave;
    // first start in [w0, w1): none means no group starts here
    uint32_t r0 = e0;
    for (uint32_t w = 0; w * kWave < e0; w++)
        if (s_st[w]) {
            r0 = w * kWave + (uint32_t)__builtin_ctzll(s_st[w]);
            break;
        }
    if (r0 >= e0) return;  // uniform: every lane read the same words
    // first start in [w1, lim), from the margin, else from memory (a group
    // that runs past the margin); lim when there is none
    uint32_t r1 = nld;
    for (uint32_t w = e0 / kWave; w < nwl; w++) {
        uint64_t b = s_st[w];
        if (w == e0 / kWave) b &= ~0ull << (e0 % kWave);
        if (b) {
            r1 = w * kWave + (uint32_t)__builtin_ctzll(b);
            break;
        }
    }
    if (r1 >= nld && w0 + nld < lim) {
        for (uint64_t c0 = w0 + nld; c0 < lim; c0 += kLThreads) {
            const uint64_t i = c0 + threadIdx.x;
            const bool st = i < lim && group_start(L, i);
            const uint64_t m = __ballot(st);
            if (lane == 0 && m) atomicMin(&s_r1, (unsigned long long)(c0 + wave * kWave + (uint64_t)__builtin_ctzll(m)));
            if (__syncthreads_or(st)) break;
        }
        __syncthreads();
        r1 = (uint32_t)(s_r1 - w0);
        if (s_r1 == lim && lim < L.n) {  // the last group of the window runs past w1 + kLMax: hand it over
            uint32_t last = r0;
            for (uint32_t w = 0; w * kWave < e0; w++) {
                uint64_t b = s_st[w];
                if ((w + 1) * kWave > e0) b &= e0 % kWave ? (1ull << (e0 % kWave)) - 1 : ~0ull;
                if (b) last = w * kWave + 63 - (uint32_t)__builtin_clzll(b);
            }
            r1 = last;
            if (threadIdx.x == 0) L.ovf[atomicAdd(L.novf, 1ull)] = w0 + last;
        } else {  // the rest of the window's last group, and no start inside it
            for (uint32_t j = nld + threadIdx.x; j < r1; j += kLThreads) s_key[j] = L.keys[w0 + j];
            for (uint32_t w = nwl + threadIdx.x; w * kWave < r1; w += kLThreads) s_st[w] = 0;
            __syncthreads();
        }
    }
    // groups [r0, r1): each tuple's rank inside its group by (low bits, position)
    const uint32_t m = r1 - r0;
    constexpr int kQ = kLCap / kLThreads;
    uint64_t kr[kQ];  // keys; the order keys (low bits << 12) | index go to LDS
    const uint64_t lowmask = (1ull << L.shift) - 1;
#pragma unroll
    for (int q = 0; q < kQ; q++) {
        const uint32_t j = q * kLThreads + threadIdx.x;
        if (j < m) kr[q] = s_key[r0 + j];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kQ; q++) {
        const uint32_t j = q * kLThreads + threadIdx.x;
        if (j < m) s_key[r0 + j] = ((kr[q] & lowmask) << 12) | j;
    }
    __syncthreads();
    // The pass runs in place: a tuple already at its rank (every tuple of a
    // group that holds one k-mer) is not written, and a tuple that moves reads
    // its label and adj before the barrier and is written after it.  The
    // workgroup owns its groups' positions; its neighbours read keys there
    // only for their high bits, which a permutation inside a group keeps.
    uint32_t rk[kQ], lb[kQ], ab[kQ];  // rank (~0: stays), label, adj: the loads wait for nothing before the writes
#pragma unroll
    for (int q = 0; q < kQ; q++) rk[q] = ~0u;
#pragma unroll
    for (int q = 0; q < kQ; q++) {
        const uint32_t j = q * kLThreads + threadIdx.x;
        if (q * kLThreads >= m) break;
        if (j >= m) continue;
        const uint32_t p = r0 + j;  // position relative to w0
        uint32_t w = p / kWave;
        const uint32_t bit = p % kWave;
        const uint64_t upto = bit == 63 ? ~0ull : (2ull << bit) - 1;
        uint64_t bs = s_st[w] & upto;  // r0 is a start: the search ends
        while (!bs) bs = s_st[--w];
        const uint32_t gs = w * kWave + 63 - (uint32_t)__builtin_clzll(bs);
        w = p / kWave;
        uint64_t be = s_st[w] & ~upto;
        while (!be && ++w * kWave < r1) be = s_st[w];
        const uint32_t ge = be ? min(r1, w * kWave + (uint32_t)__builtin_ctzll(be)) : r1;
        const uint64_t cj = s_key[p];
        // rank = gs + the group's order keys below cj, 4 keys a step (one LDS
        // round trip for 4; the keys past the group's end are masked, so no
        // remainder loop: 13.9 ms against 14.2 one key a step)
        uint32_t rank = gs;
        for (uint32_t x = gs; x < ge; x += 4) {
            const uint64_t k0 = s_key[x], k1 = s_key[x + 1], k2 = s_key[x + 2], k3 = s_key[x + 3];
            rank += (uint32_t)(k0 < cj) + (uint32_t)(x + 1 < ge && k1 < cj) + (uint32_t)(x + 2 < ge && k2 < cj) +
                    (uint32_t)(x + 3 < ge && k3 < cj);
        }
        if (rank != p) {
            rk[q] = rank;
            ab[q] = L.adj[w0 + p];
            lb[q] = L.lab[w0 + p];
        }
    }
    __syncthreads();
    uint32_t nmoved = 0;
#pragma unroll
    for (int q = 0; q < kQ; q++) {
        if (rk[q] == ~0u) continue;
        const uint64_t dst = w0 + rk[q];
        L.keys[dst] = kr[q];
        L.lab[dst] = lb[q];
        L.adj[dst] = (uint8_t)ab[q];
        nmoved++;
    }
    // one plain store per wave (no barrier, no atomic: millions of adds on a
    // few addresses serialise)
    for (int d = kWave / 2; d > 0; d >>= 1) nmoved += __shfl_down(nmoved, d, kWave);
    if (lane == 0) L.moved[(uint64_t)blockIdx.x * (kLThreads / kWave) + wave] = nmoved;
}

// sum of n counts into *out (grid-stride, one atomic per workgroup)
__global__ __launch_bounds__(256) void sum_u32_kernel(const uint32_t *__restrict__ v, uint64_t n,
                                                      unsigned long long *__restrict__ out)
{
    uint64_t t = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) t += v[i];
    for (int d = kWave / 2; d > 0; d >>= 1) t += __shfl_down(t, d, kWave);
    __shared__ uint64_t s[256 / kWave];
    if (__lane_id() == 0) s[threadIdx.x / kWave] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t a = 0;
        for (int w = 0; w < 256 / kWave; w++) a += s[w];
        if (a) atomicAdd(out, (unsigned long long)a);
    }
}

// end of each long group: the next group start
__global__ void long_end_kernel(const uint64_t *__restrict__ keys, uint64_t n, int shift,
                                const uint64_t *__restrict__ starts, uint64_t nl, const uint64_t *__restrict__ seg,
                                int nseg, uint64_t *__restrict__ ends)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nl) return;
    const uint64_t s = starts[g], hk = keys[s] >> shift;
    // first index > s with a different high part (sorted by it), at most the next sample's start
    uint64_t lo = s + 1, hi = nseg ? next_seg_start(seg, nseg, s + 1) : n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if ((keys[mid] >> shift) == hk) lo = mid + 1; else hi = mid;
    }
    ends[g] = lo;
}

// Low bits ranked locally: the number of low key bits, 0 for the plain
// all-global order.  About 30 high bits keep groups small; K = 32 keeps the
// global order (the reference's forward code is 0 after the first base, so
// most canonical k-mers are equal there).
static int local_bits(int K)
{
    const int bits = 2 * K;
    if (K >= 32 || bits < 40) return 0;
    return ((bits - 30) / 8) * 8;
}

static int sort_local_bits(int K) { return getenv("MP_SORT_GLOBAL") ? 0 : local_bits(K); }

KeyLayout byte_layout()
{
    KeyLayout L;
    for (int b = 0; b < 8; b++) {
        L.shift[b] = (uint8_t)(8 * b);
        L.bits[b] = 8;
    }
    return L;
}

// The scans' layout: bytes, except that r = 2K - 48 bits above bit 48 with
// 8 < r < 16 go evenly to the passes of bytes 6 and 7 (those passes run only
// on bits the scans sort globally, above the local pass's low bits).
KeyLayout scan_layout(const mp_ctx *ctx)
{
    KeyLayout L = byte_layout();
    const int r = 2 * ctx->cfg.k - 48;
    int bits = ctx->cfg.sort_digit_bits;
    if (const char *e = getenv("MP_SORT_DIGITS")) bits = atoi(e);  // tests run both layouts
    if (bits != 8 && r > 8 && r < 16 && sort_local_bits(ctx->cfg.k) <= 48) {
        const int w = (r + 1) / 2;
        L.bits[6] = (uint8_t)w;
        L.shift[7] = (uint8_t)(48 + w);
        L.bits[7] = (uint8_t)(r - w);
    }
    return L;
}

uint32_t scan_pass_mask(int K)
{
    const int S = sort_local_bits(K);
    uint32_t mask = 0;
    for (int b = S / 8; b < 8; b++) mask |= 1u << (b + 1);
    return mask;
}

int radix_sort_tuples(mp_ctx *ctx, uint64_t n, int *passes_out, const GenSource *gen, const SortSegments *segs)
{
    // k-mer bytes only: within a run nothing downstream depends on the adj
    // order except the quirk's first tuples and the tuple dump, which derive it
    const int S = sort_local_bits(ctx->cfg.k);
    const uint32_t mask = scan_pass_mask(ctx->cfg.k);
    const KeyLayout KL = scan_layout(ctx);
    MP_TRY(radix_sort_generic(ctx, ctx->keys, ctx->lab, ctx->adj, &ctx->cur, n, mask, passes_out, nullptr, 0,
                              nullptr, gen, segs, &KL));
    if (S == 0 || n <= 1) return 0;
    const uint64_t wins = (n + kLWin - 1) / kLWin;
    // scratch: long-group count, moved-tuple total, long-group starts (one per window at most),
    // then the per-wave moved counts
    const uint64_t nwave = wins * (kLThreads / kWave);
    MP_TRY(ensure(ctx, ctx->scratch, (wins + 3) * sizeof(uint64_t) + nwave * sizeof(uint32_t) + 64));
    unsigned long long *novf = (unsigned long long *)ctx->scratch.p;
    uint32_t *d_moved = (uint32_t *)((uint64_t *)ctx->scratch.p + wins + 3);
    MP_TRY(fill_u32(ctx, novf, 0, 4));
    MP_TRY(fill_u32(ctx, d_moved, 0, nwave));
    const int src = ctx->cur;  // the local pass sorts this buffer in place
    LocalPass L;
    L.keys = (uint64_t *)ctx->keys[src].p;
    L.lab = (uint32_t *)ctx->lab[src].p;
    L.adj = (uint8_t *)ctx->adj[src].p;
    L.n = n;
    L.shift = S;
    L.ovf = (uint64_t *)ctx->scratch.p + 2;
    L.novf = novf;
    L.moved = d_moved;
    L.seg = nullptr;
    L.nseg = 0;
    if (!grid_fits(wins, kLThreads))  // checked before the segment table is allocated
        return ctx->fail(MP_EINVAL, "local pass of %llu windows exceeds one dispatch: use more scans",
                         (unsigned long long)wins);
    uint64_t *d_lseg = nullptr;
    if (segs && segs->ns > 0) {
        MP_TRY(tmp_alloc(ctx, (void **)&d_lseg, (segs->ns + 1) * sizeof(uint64_t)));
        MP_HIP(hipMemcpyAsync(d_lseg, segs->tpos, (segs->ns + 1) * sizeof(uint64_t), hipMemcpyHostToDevice,
                              ctx->stream));
        L.seg = d_lseg;
        L.nseg = segs->ns;
    }
    hipEventRecord(ctx->pev[22], ctx->stream);
    hipLaunchKernelGGL(local_rank_kernel, dim3((unsigned)wins), dim3(kLThreads), 0, ctx->stream, L);
    MP_HIP(hipGetLastError());
    hipEventRecord(ctx->pev[23], ctx->stream);
    hipLaunchKernelGGL(sum_u32_kernel, dim3(stride_grid(nwave)), dim3(256), 0, ctx->stream, (const uint32_t *)d_moved,
                       nwave, novf + 1);
    MP_HIP(hipGetLastError());
    uint64_t cnt[2] = {0, 0};
    MP_HIP(hipMemcpyAsync(cnt, novf, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    const uint64_t nl = cnt[0];
    ctx->local_moved = cnt[1];
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->pev[22], ctx->pev[23]);
    ctx->local_ms = ms;
    ctx->local_long = nl;
    if (nl) {  // long groups: their own low-byte passes over [start, end)
        std::vector<uint64_t> se(2 * nl);
        uint64_t *d_ends = nullptr;
        MP_HIP(hipMallocAsync((void **)&d_ends, nl * sizeof(uint64_t), ctx->stream));
        hipLaunchKernelGGL(long_end_kernel, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, ctx->stream,
                           (const uint64_t *)L.keys, n, S, (const uint64_t *)L.ovf, nl, (const uint64_t *)L.seg, L.nseg,
                           d_ends);
        MP_HIP(hipGetLastError());
        MP_HIP(hipMemcpyAsync(se.data(), L.ovf, nl * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
        MP_HIP(hipMemcpyAsync(se.data() + nl, d_ends, nl * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
        MP_HIP(hipFreeAsync(d_ends, ctx->stream));
        MP_HIP(hipStreamSynchronize(ctx->stream));
        uint32_t lmask = 0;
        for (int b = 0; b < S / 8; b++) lmask |= 1u << (b + 1);
        const float keep_avg = ctx->sort_pass_avg, keep_hist = ctx->sort_hist_ms,
                    keep_gen = ctx->sort_gen_ms, keep_count = ctx->sort_count_ms;  // stats of the main passes
        const int keep_graph = ctx->sort_graph_last;
        for (uint64_t g = 0; g < nl; g++) {
            const uint64_t a = se[g], e = se[nl + g];
            DevArray kv[2], lv[2], av[2];
            for (int b = 0; b < 2; b++) {
                kv[b].p = (uint64_t *)ctx->keys[b].p + a;
                lv[b].p = (uint32_t *)ctx->lab[b].p + a;
                av[b].p = (uint8_t *)ctx->adj[b].p + a;
                kv[b].bytes = (e - a) * 8;
                lv[b].bytes = (e - a) * 4;
                av[b].bytes = e - a;
            }
            int c = src, np = 0;
            MP_TRY(radix_sort_generic(ctx, kv, lv, av, &c, e - a, lmask, &np, nullptr, 0, nullptr));
            if (c != src) {  // an odd number of passes ran: the range is in the other buffer
                MP_HIP(hipMemcpyAsync(kv[src].p, kv[c].p, (e - a) * 8, hipMemcpyDeviceToDevice, ctx->stream));
                MP_HIP(hipMemcpyAsync(lv[src].p, lv[c].p, (e - a) * 4, hipMemcpyDeviceToDevice, ctx->stream));
                MP_HIP(hipMemcpyAsync(av[src].p, av[c].p, e - a, hipMemcpyDeviceToDevice, ctx->stream));
            }
        }
        MP_HIP(hipStreamSynchronize(ctx->stream));
        ctx->sort_pass_avg = keep_avg;
        ctx->sort_count_ms = keep_count;
        ctx->sort_graph_last = keep_graph;
        ctx->sort_hist_ms = keep_hist;
        ctx->sort_gen_ms = keep_gen;
    }
    if (d_lseg) MP_TRY(tmp_free(ctx, d_lseg));
    *passes_out += 1;
    return 0;
}

}  // namespace mp
