// Frequency filter, union-find and the post-scan partition plan.
//
// State per read (32-bit ids): parent[i], bad[i] (the reference's BAD
// sentinel 0xBADBADBAD in parent), filtered[i] (filteredReads).  A BAD read
// keeps parent[i] == i: in the reference a label is only ever marked BAD while
// it is a root, and findRoot stops at a node whose parent is BAD, which is the
// same as stopping at a self-parent (unionfind.h:19-55).
//
// Filter (markMischiefKmers, metapart.c:2270-2519): tiles of tuples staged in
// LDS; the lane holding a run's first tuple evaluates the run unless it is
// the last run of its segment.  Zeroing is recorded in a bitmap (the keys stay
// as sorted; the union and the quirk read the bitmap, the tuple dump applies it).
//
// Union (unionFindPar2 + unionFind, metapart.c:2520-2800): the reference
// star-unions the non-BAD labels of each run and always hooks the smaller root
// under the larger one, so every component's root is its largest id.  Here
// each run start does the same with a lock-free union-find: find with path
// halving, hook by compare-and-swap of the smaller root's parent.  Since
// parent[x] >= x always holds, the final roots are the component maxima
// whatever order the lanes run in.
#include "mp_internal.h"

#include <functional>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

namespace mp {

struct GParams {
    const uint64_t *keys;
    const uint8_t *adj;
    const uint32_t *lab;
    uint64_t n;
    const uint64_t *seg;  // T+1 tuple positions
    int T;
    int open_last;        // the last segment goes on past these tuples (a part of a two-sweep scan)
    uint8_t *bad;
    uint8_t *filtered;
    uint32_t *zbits;
    uint32_t *parent;
    unsigned long long *stats;  // [0] evaluated runs, [1] zeroed runs
};

__global__ void seg_pos_kernel(const uint64_t *__restrict__ keys, uint64_t n, int bshift,
                               const uint64_t *__restrict__ bounds, int T, uint64_t *__restrict__ pos)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > T) return;
    const uint64_t target = bounds[t];
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if ((keys[mid] >> bshift) < target) lo = mid + 1; else hi = mid;
    }
    pos[t] = lo;
}

// A run's verdict from its in/out neighbour-class tallies (metapart.c:2300-2359,
// 2419): 0 = nothing, kVZero = zero the k-mer (both sides branch), else
// kVMark with the rule each tuple's label is filtered by.
constexpr uint32_t kVMark = 1, kVZero = 2, kVRatio = 4, kVIn = 8, kVOut = 16;

// One side of metapart.c:2300-2333 over classes 0-6: the number of classes
// present, the largest count and its class (the first on ties), and the
// second-largest count (with multiplicity).  Branch-free: the maximum of
// count << 3 | (7 - class) is the largest count at its lowest class.
__device__ __forceinline__ void class_stats(const uint32_t cnt[8], int &n, uint32_t &mx, int &arg, uint32_t &second)
{
    uint32_t best = 0;
    n = 0;
#pragma unroll
    for (int c = 0; c < 7; c++) {
        best = max(best, cnt[c] ? (cnt[c] << 3) | (uint32_t)(7 - c) : 0u);
        n += cnt[c] != 0;
    }
    mx = best >> 3;
    arg = 7 - (int)(best & 7);
    second = 0;
#pragma unroll
    for (int c = 0; c < 7; c++) second = max(second, c == arg ? 0u : cnt[c]);
}

__device__ __forceinline__ uint32_t run_verdict(const uint32_t in[8], const uint32_t out[8], uint64_t len)
{
    int nIn, nOut, maxIn, maxOut;
    uint32_t mIn, mOut, sIn, sOut;
    class_stats(in, nIn, mIn, maxIn, sIn);
    class_stats(out, nOut, mOut, maxOut, sOut);
    if (!(nIn > 1 || nOut > 1)) return 0;
    if (nIn > 1 && nOut > 1) return kVZero;
    // (float)smax / degree compared with the double 0.2 (metapart.c:2356-2359, 2419)
    const float deg = (float)(int)len;
    const bool ratio = (double)__fdiv_rn((float)sIn, deg) > 0.2 || (double)__fdiv_rn((float)sOut, deg) > 0.2;
    return kVMark | (ratio ? kVRatio : 0) | (nIn > 1 ? kVIn : 0) | (nOut > 1 ? kVOut : 0) | ((uint32_t)maxIn << 8) |
           ((uint32_t)maxOut << 12);
}

// filteredReads for a tuple of a kVMark run; the refilter step turns it into
// BAD (the label is a root here, so that step's parent reset changes nothing for it)
__device__ __forceinline__ bool tuple_filtered(uint32_t v, uint32_t a)
{
    return ((v & kVIn) && (a & 7) != ((v >> 8) & 7)) || ((v & kVOut) && ((a >> 3) & 7) != ((v >> 12) & 7)) ||
           (v & kVRatio);
}

__device__ void zero_run(const GParams &G, uint64_t i, uint64_t j)
{
    for (uint64_t t = i; t < j;) {
        const uint64_t w = t >> 5;
        const uint64_t e = ((w + 1) << 5) < j ? ((w + 1) << 5) : j;
        const uint32_t nb = (uint32_t)(e - t);
        const uint32_t m = (nb == 32 ? 0xFFFFFFFFu : ((1u << nb) - 1)) << (t & 31);
        atomicOr(&G.zbits[w], m);
        t = e;
    }
}

// Evaluates run [i, j) (not the last run of its segment) from memory and
// applies its verdict.  Returns 1, or 2 when zeroed.
__device__ int eval_run_mem(const GParams &G, uint64_t i, uint64_t j)
{
    uint32_t in[8], out[8];
#pragma unroll
    for (int c = 0; c < 8; c++) in[c] = out[c] = 0;
    for (uint64_t t = i; t < j; t++) {
        const uint32_t a = G.adj[t];
        in[a & 7]++;
        out[(a >> 3) & 7]++;
    }
    const uint32_t v = run_verdict(in, out, j - i);
    if (v & kVZero) {
        zero_run(G, i, j);
        return 2;
    }
    if (v & kVMark)
        for (uint64_t t = i; t < j; t++)
            if (tuple_filtered(v, G.adj[t])) G.filtered[G.lab[t]] = 1;
    return 1;
}

// Filter over tiles of 6144 tuples, looking 256 tuples past the tile so that
// a run crossing its end is usually whole.  Of the tuples only bitmaps are
// staged in LDS: lane l of the workgroup owns tuples l, l+256, ... (a wave
// owns 64 consecutive tuples, a "word").  Three steps:
//   1. per word, wave ballots give the run starts (a key change, or a segment
//      start from the tile's segment bitmap) and six bit planes of the adj
//      byte (three each of the in and out neighbour classes); a run of one
//      tuple (the next tuple of the word starts a run too) has no verdict,
//      and the wave's other run starts go to the wave's list;
//   2. one lane per listed run finds the run's end in the start bitmap, its
//      class tallies by popcounts over the run's words, and its verdict; a run
//      longer than the tuples looked at is evaluated from memory;
//   3. every tuple whose run starts in the tile applies the run's verdict to
//      its label (the tile before owns the runs that start before).
// Besides the filter's effects it records two bitmaps for the union: every
// run start and the starts of the runs the union must join (k-mer not zero
// and not zeroed, two tuples or more).  Word positions, masks and list
// counts are wave-uniform and stay in scalar registers, and no step searches
// per tuple.  The tile's fixed cost (its barriers, the segment bitmap, the
// label gather, the counters) is what bounds the kernel, so the tile is as
// large as keeps the workgroups a CU holds (filter time at 20 M reads,
// k = 31: 2048 tuples a tile 23.0 ms, 4096 12.5, 6144 10.9, 8192 11.8).
constexpr int kFLanes = 256, kFPer = 24, kFTile = kFLanes * kFPer, kFHalo = 256, kFStage = kFTile + kFHalo,
              kFIter = kFStage / kFLanes, kFWords = kFStage / 64, kFWaves = kFLanes / 64,
              kFList = 33 * kFPer;  // a wave's listed runs: at most 33 a word (32 of two tuples or more, and the last)
static_assert(kFWaves == 4, "the run lists assume 4 waves");
static_assert(kFIter <= 64, "one bit a word in the marking step's mask");

// the number of bits of m below this lane
__device__ __forceinline__ uint32_t bits_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// tuples of class C (0-6) among the ones in a (the mask with plane 0's bit
// of C applied): a & (plane 1 or its complement) & (plane 2 or its
// complement), one three-input bit operation a 32-bit half (its truth table
// holds the one row a = 1, p1 = bit 1 of C, p2 = bit 2 of C)
template <int C>
__device__ __forceinline__ uint32_t class_tally(uint64_t a, uint64_t p1, uint64_t p2)
{
    constexpr unsigned kRow = 4 | ((C & 2) ? 2 : 0) | ((C & 4) ? 1 : 0);
    return (uint32_t)__builtin_popcount(__builtin_amdgcn_bitop3_b32((uint32_t)a, (uint32_t)p1, (uint32_t)p2, 1u << kRow)) +
           (uint32_t)__builtin_popcount(__builtin_amdgcn_bitop3_b32((uint32_t)(a >> 32), (uint32_t)(p1 >> 32),
                                                                    (uint32_t)(p2 >> 32), 1u << kRow));
}

// tallies of the tuples in m by class (0-6) from the in side's three bit
// planes (class bit b in plane b) and the out side's (planes 3-5): class c's
// in tally in the low half of io[c], its out tally in the high half (a run
// here holds at most kFStage < 2^16 tuples; half the registers keep the run
// loop from spilling)
__device__ __forceinline__ void tally_classes(const uint64_t *pl, uint64_t m, uint32_t io[7])
{
#pragma unroll
    for (int s = 0; s < 2; s++) {  // one side at a time: fewer masks live
        const uint64_t p0 = pl[3 * s], p1 = pl[3 * s + 1], p2 = pl[3 * s + 2];
        const uint64_t a0 = m & ~p0, a1 = m & p0;
        io[0] += class_tally<0>(a0, p1, p2) << (16 * s);
        io[1] += class_tally<1>(a1, p1, p2) << (16 * s);
        io[2] += class_tally<2>(a0, p1, p2) << (16 * s);
        io[3] += class_tally<3>(a1, p1, p2) << (16 * s);
        io[4] += class_tally<4>(a0, p1, p2) << (16 * s);
        io[5] += class_tally<5>(a1, p1, p2) << (16 * s);
        io[6] += class_tally<6>(a0, p1, p2) << (16 * s);
    }
}

// run_verdict of a staged run (len <= kFStage) in the form step 3 applies:
// bit 0 marks (kVMark), bit 1 the out side (kVOut; else kVIn), bit 2 every
// tuple (kVRatio), bits 3-5 the side's largest class, which a tuple of the
// side's other classes is filtered for; 0 = nothing to mark.  zero: both
// sides branch (kVZero).  io: tally_classes' packed tallies.
__device__ __forceinline__ uint32_t staged_verdict(const uint32_t io[7], uint32_t len, bool &zero)
{
    uint32_t nIn = 0, nOut = 0;
#pragma unroll
    for (int c = 0; c < 7; c++) {
        nIn += (io[c] & 0xFFFF) != 0;
        nOut += (io[c] >> 16) != 0;
    }
    zero = nIn > 1 && nOut > 1;
    if (zero || (nIn <= 1 && nOut <= 1)) return 0;
    // one side branches: the other's second-largest count is 0, so only this
    // side decides the ratio.  The two largest of count << 3 | (7 - class):
    // the largest count at its lowest class (two classes are present, so it
    // is not an empty one), and the second-largest count with multiplicity.
    const bool side_in = nIn > 1;
    uint32_t b1 = 0, b2 = 0;
#pragma unroll
    for (int c = 0; c < 7; c++) {
        const uint32_t v = (((io[c] >> (side_in ? 0 : 16)) & 0xFFFF) << 3) | (uint32_t)(7 - c);
        b2 = max(b2, min(b1, v));
        b1 = max(b1, v);
    }
    const uint32_t arg = 7 - (b1 & 7), second = b2 >> 3;
    // (float)second / (float)len > 0.2 in double (metapart.c:2356-2359,
    // 2419): the rounded quotient exceeds 0.2 exactly when it is at least
    // 0x3E4CCCCD, i.e. when second / len lies above the midpoint
    // 26843545 * 2^-27 below it (a tie rounds to the even 0x3E4CCCCC); both
    // operands are below 2^24, so the float conversions are exact
    const bool ratio = ((uint64_t)second << 27) > 26843545ull * len;
    return 1u | (side_in ? 0u : 2u) | (ratio ? 4u : 0u) | (arg << 3);
}

__global__ __launch_bounds__(kFLanes, 8) void filter_tile_kernel(GParams G, uint64_t *__restrict__ rs_bits,
                                                                 uint64_t *__restrict__ st_bits)
{
    __shared__ uint64_t s_st[kFWords], s_rs[kFWords], s_sg[kFWords];  // run starts, union runs, segment starts
    __shared__ uint64_t s_pl[kFWords][6];                 // bits 0-2 of the in class, then of the out class
    __shared__ int32_t s_ls[kFWords];                     // the last run start up to each word's end, -1: none
    __shared__ __align__(16) uint8_t s_vd[kFTile];        // verdicts (staged_verdict) at the listed run starts
    __shared__ uint16_t s_list[kFWaves][kFList];          // each wave's listed run starts
    __shared__ uint32_t s_nl[kFWaves], s_anyseg;
    __shared__ unsigned long long s_tail;  // the first segment boundary at or past the tuples looked at
    const uint64_t t0 = (uint64_t)blockIdx.x * kFTile;
    const uint32_t cnt = (uint32_t)min((uint64_t)kFTile, G.n - t0);    // run starts owned
    const uint32_t cnt_s = (uint32_t)min((uint64_t)kFStage, G.n - t0);  // tuples looked at
    const uint32_t tid = threadIdx.x, lane = tid % kWave;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    if (tid < kFWords) {
        s_rs[tid] = 0;
        s_sg[tid] = 0;
    }
    if (tid == 0) {
        s_tail = ~0ull;
        s_anyseg = 0;
    }
    static_assert(kFTile % 16 == 0, "s_vd cleared 16 bytes a store");
    for (uint32_t x = tid; x < kFTile / 16; x += kFLanes) reinterpret_cast<uint4 *>(s_vd)[x] = make_uint4(0u, 0u, 0u, 0u);
    // each word's key, previous key and adj byte are loaded one word ahead
    uint64_t nk = 0, nkp = 0;
    uint32_t na = 0;
    auto load = [&](int q) {
        const uint32_t j = q * kFLanes + tid;
        if (j < cnt_s) {
            const uint64_t i = t0 + j;
            nk = G.keys[i];
            nkp = i ? G.keys[i - 1] : 0;
            na = G.adj[i];
        }
    };
    // the segment starts among the tuples looked at (each starts a run; the
    // filter evaluates no run that ends its segment) and the first boundary
    // past them.  The positions run from seg[0] = 0 to seg[T] = n; the first
    // 256 are read with the first word's tuples, before the barrier.
    const uint64_t sp0 = tid <= (uint32_t)G.T ? G.seg[tid] : ~0ull;
    load(0);
    __syncthreads();
    auto note_seg = [&](uint32_t t, uint64_t p) {
        if (t < (uint32_t)G.T && p >= t0 && p < t0 + cnt_s) {
            atomicOr((unsigned long long *)&s_sg[(p - t0) / kWave], 1ull << ((p - t0) % kWave));
            s_anyseg = 1;
        }
        if (p >= t0 + cnt_s && p != ~0ull) atomicMin(&s_tail, (unsigned long long)p);
    };
    note_seg(tid, sp0);
    for (uint32_t t = tid + kFLanes; t <= (uint32_t)G.T; t += kFLanes) note_seg(t, G.seg[t]);
    __syncthreads();
    const bool anyseg = s_anyseg != 0;
    uint32_t ev = 0, z = 0;     // runs evaluated (the last run of a segment is not), runs zeroed
    uint32_t nl = 0, evs = 0;   // the wave's listed runs, its evaluated runs of one tuple
#pragma unroll
    for (int q = 0; q < kFIter; q++) {
        const uint32_t jw = q * kFLanes + wave * kWave;  // the word's first tuple
        const uint64_t k = nk, kp = nkp;
        const uint32_t an = na;
        if (q + 1 < kFIter) load(q + 1);
        if (jw >= cnt_s) break;
        const uint32_t j = jw + lane;
        const bool valid = j < cnt_s;
        const uint32_t a = valid ? an : 0u;
        uint64_t sg = 0;  // uniform: read into scalar registers, so that the masks below are scalar work
        if (anyseg) {
            // readfirstlane returns a signed int: each half goes through uint32_t,
            // or bit 31 of the low half would be copied into bits 32-63
            const uint64_t v = s_sg[jw / kWave];
            const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)),
                           lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
            sg = ((uint64_t)hi << 32) | lo;
        }
        const uint64_t mst = __builtin_amdgcn_ballot_w64(valid && kp != k) | sg;
        uint64_t pl[6];
#pragma unroll
        for (int b = 0; b < 6; b++) pl[b] = __builtin_amdgcn_ballot_w64((a >> b) & 1);  // a = 0 past the tuples
        const uint64_t own = jw + 64 <= cnt ? ~0ull : (jw < cnt ? (1ull << (cnt - jw)) - 1 : 0);
        // runs of one tuple: the tuple after it (same word) starts a run.  Its
        // tallies hold one class a side, so no verdict; it is evaluated unless
        // the next tuple starts a segment (it cannot end the tuples: the next
        // one is valid)
        const uint64_t msingle = mst & (mst >> 1) & own, mlist = mst & own & ~msingle;
        evs += (uint32_t)__popcll(msingle & ~(sg >> 1));
        if ((mlist >> lane) & 1) s_list[wave][nl + bits_below(mlist)] = (uint16_t)j;
        nl += (uint32_t)__popcll(mlist);
        if (lane == 0) {
            const uint32_t wd = jw / kWave;
            s_st[wd] = mst;
#pragma unroll
            for (int b = 0; b < 6; b++) s_pl[wd][b] = pl[b];
            if (jw < cnt) st_bits[(t0 + jw) / 64] = mst;
        }
    }
    if (lane == 0) s_nl[wave] = nl;
    __syncthreads();
    const uint32_t nw = (cnt_s + kWave - 1) / kWave;
    if (wave == 0) {  // s_ls: a max-scan of each word's last start, 64 words at a time
        int carry = -1;
        for (uint32_t w0 = 0; w0 < (uint32_t)kFWords; w0 += kWave) {
            const uint32_t w = w0 + lane;
            const uint64_t v = w < nw ? s_st[w] : 0;
            int ls = v ? (int)(w * kWave + 63 - (uint32_t)__builtin_clzll(v)) : carry;
#pragma unroll
            for (int d = 1; d < kWave; d <<= 1) {
                const int o = __shfl_up(ls, d, kWave);
                if (lane >= (uint32_t)d) ls = max(ls, o);
            }
            if (w < (uint32_t)kFWords) s_ls[w] = ls;
            carry = __shfl(ls, kWave - 1, kWave);
        }
    }
    const uint32_t n0 = s_nl[0], n1 = n0 + s_nl[1], n2 = n1 + s_nl[2], nruns = n2 + s_nl[3];
    const uint64_t tail = s_tail;
    for (uint32_t r = tid; r < nruns; r += kFLanes) {
        const uint32_t lw = (r >= n0) + (r >= n1) + (r >= n2);
        const uint32_t j = s_list[lw][r - (lw == 0 ? 0 : lw == 1 ? n0 : lw == 2 ? n1 : n2)];
        const uint64_t i = t0 + j;
        // the run's k-mer is zero only if the run begins its segment (the
        // segment's tuples are sorted, 0 is the smallest k-mer): read from
        // memory then, or for a run that goes on past the tuples looked at
        const bool seg_start = anyseg && ((s_sg[j / kWave] >> (j % kWave)) & 1);
        // the run ends at the next start among the tuples looked at, else it continues past them
        uint32_t w = j / kWave;
        uint64_t b = (j % kWave) == 63 ? 0 : s_st[w] & (~0ull << (j % kWave + 1));
        while (!b && ++w < nw) b = s_st[w];
        const uint32_t je = b ? w * kWave + (uint32_t)__builtin_ctzll(b) : cnt_s;
        uint64_t e = t0 + je;
        bool evald;  // the last run of a segment is never evaluated
        if (je < cnt_s) {
            // its segment ends at the next segment start, which is no earlier than its end
            evald = !(anyseg && ((s_sg[je / kWave] >> (je % kWave)) & 1));
        } else {
            // no run, so no segment, starts after it here: its segment ends at the first boundary past
            const uint64_t k = G.keys[i];
            while (e < tail && G.keys[e] == k) e++;
            evald = e != tail || (G.open_last && e == G.n);
        }
        bool zeroed = false;
        uint32_t c = 0;
        if (evald) {
            ev++;
            if (e <= t0 + cnt_s) {
                uint32_t io[7];
#pragma unroll
                for (int x = 0; x < 7; x++) io[x] = 0;
                for (uint32_t wd = j / kWave; wd * kWave < je; wd++) {
                    const uint32_t lo = max(j, wd * kWave) - wd * kWave, hi = min(je, (wd + 1) * kWave) - wd * kWave;
                    const uint64_t m = (hi == 64 ? ~0ull : ((1ull << hi) - 1)) & ~((1ull << lo) - 1);
                    tally_classes(s_pl[wd], m, io);
                }
                c = staged_verdict(io, (uint32_t)(e - i), zeroed);
                if (zeroed) zero_run(G, i, e);
            } else {  // longer than the tuples looked at: memory (marks applied there)
                zeroed = eval_run_mem(G, i, e) == 2;
            }
            z += zeroed;
        }
        s_vd[j] = (uint8_t)c;
        if (!zeroed && e - i >= 2 && !(seg_start && G.keys[i] == 0))
            atomicOr((unsigned long long *)&s_rs[j / kWave], 1ull << (j % kWave));
    }
    __syncthreads();
    if (tid * 64 < cnt) rs_bits[t0 / 64 + tid] = s_rs[tid];
    // each tuple's run start: the last start at or before it in its word, else
    // the last one before the word (none: an earlier tile owns the run); its
    // adj bits from the word's bit planes.  The lane's filtered tuples first
    // (bit q), then their labels loaded 8 at a time, then their marks: one
    // load latency for 8 tuples
    const uint64_t upto = lane == 63 ? ~0ull : (2ull << lane) - 1;
    uint64_t fm = 0;
#pragma unroll
    for (int q = 0; q < kFIter; q++) {
        const uint32_t jw = q * kFLanes + wave * kWave;
        if (jw >= cnt_s) break;
        const uint32_t j = jw + lane;
        uint32_t a = 0;
#pragma unroll
        for (int x = 0; x < 6; x++) a |= (uint32_t)((s_pl[jw / kWave][x] >> lane) & 1) << x;
        const uint64_t b = s_st[jw / kWave] & upto;
        const int prev = jw ? s_ls[jw / kWave - 1] : -1;
        const int sj = b ? (int)(jw + 63 - (uint32_t)__builtin_clzll(b)) : prev;
        const bool ok = j < cnt_s && sj >= 0 && sj < (int)cnt;
        const uint32_t c = s_vd[ok ? sj : 0];  // no branch: the LDS reads issue together
        const uint32_t f = (c & 1) & ((c >> 2) | (uint32_t)(((a >> (3 * ((c >> 1) & 1))) & 7) != (c >> 3)));
        fm |= (uint64_t)(ok ? f : 0u) << q;
    }
    constexpr int kFChunk = 8;
#pragma unroll
    for (int q0 = 0; q0 < kFIter; q0 += kFChunk) {
        uint32_t lv[kFChunk];
#pragma unroll
        for (int u = 0; u < kFChunk; u++)
            if (q0 + u < kFIter && ((fm >> (q0 + u)) & 1)) lv[u] = G.lab[t0 + (q0 + u) * kFLanes + tid];
        // every load above completes here, once: the stores below then wait on nothing
        // (left to itself the compiler waits before each store, the previous store included)
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), expcnt and lgkmcnt unchanged (gfx9 encoding)
#pragma unroll
        for (int u = 0; u < kFChunk; u++)
            if (q0 + u < kFIter && ((fm >> (q0 + u)) & 1)) G.filtered[lv[u]] = 1;
    }
    block_add(ev + (lane == 0 ? evs : 0u), &G.stats[0]);
    block_add(z, &G.stats[1]);
}

// metapart.c:2500-2517: a read filtered in any scan is BAD again after the filter
__global__ void refilter_kernel(uint8_t *__restrict__ bad, const uint8_t *__restrict__ filtered,
                                uint32_t *__restrict__ parent, uint64_t N, unsigned long long *__restrict__ nbad)
{
    uint32_t cnt = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        if (filtered[i] && !bad[i]) {
            bad[i] = 1;
            parent[i] = (uint32_t)i;
        }
        cnt += bad[i] != 0;
    }
    block_add(cnt, nbad);
}

__device__ __forceinline__ uint32_t uf_find(uint32_t *parent, uint32_t x)
{
    for (;;) {
        const uint32_t p = ld_agent(&parent[x]);
        if (p == x) return x;
        const uint32_t gp = ld_agent(&parent[p]);
        if (gp != p) __hip_atomic_store(&parent[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = gp;
    }
}

__device__ __forceinline__ void uf_union(uint32_t *parent, uint32_t a, uint32_t b)
{
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a < b) { const uint32_t t = a; a = b; b = t; }
        // hook the smaller root under the larger (unionfind.h:58-69 with u > v)
        uint32_t expect = b;
        if (__hip_atomic_compare_exchange_strong(&parent[b], &expect, a, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

// bad[] bytes -> one bit per read, small enough to stay in L2 for the union's
// random lookups (20M reads: 2.5 MB)
__global__ void bad_bits_kernel(const uint8_t *__restrict__ bad, uint64_t N, uint64_t *__restrict__ bits)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t m = __ballot(i < N && bad[i] != 0);
    if (__lane_id() == 0 && i < N) bits[i >> 6] = m;
}

__device__ __forceinline__ bool is_bad(const uint64_t *bits, uint32_t x) { return (bits[x >> 6] >> (x & 63)) & 1; }

// One lane per tuple; a lane whose tuple starts a run to join finds the run's
// end in the run-start bitmap and star-joins the non-BAD labels to the first
// non-BAD one (metapart.c:2539-2616).
// Union of every run to join (metapart.c:2520-2800).  The reference joins a
// run's live labels to its first one (a star); a chain through the live
// labels in run order joins the same sets, and lets a lane take one tuple:
// each wave takes a 64-tuple word of the run bitmaps, loads its labels
// (coalesced), and every live tuple (in a run to join, label not BAD) joins
// the nearest live tuple before it in its run -- found by a ballot inside the
// word, or, for the first live tuple of a run that began in an earlier word,
// by a walk back through that run.  Most tuples of this workload have BAD
// labels, so the work is the label and BAD-bit reads, not the finds.
constexpr int kUW = 4;  // bitmap words a wave takes per step, their loads in flight together

__global__ __launch_bounds__(256) void union_kernel(GParams G, const uint64_t *__restrict__ rs_bits,
                                                    const uint64_t *__restrict__ st_bits,
                                                    const uint64_t *__restrict__ badbits)
{
    const uint64_t nw = (G.n + 63) / 64;
    const unsigned lane = __lane_id();
    const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / kWave;
    const uint64_t lt = lanemask_lt();
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    for (uint64_t w0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave * kUW; w0 < nw;
         w0 += nwaves * kUW) {
        // the words' bitmaps and every lane's labels, loaded together (a label
        // is used only in a joined run, but the lines are fetched anyway)
        uint64_t st[kUW], rs[kUW];
        uint32_t lb[kUW];
#pragma unroll
        for (int u = 0; u < kUW; u++) {
            st[u] = w0 + u < nw ? st_bits[w0 + u] : 0;
            rs[u] = w0 + u < nw ? rs_bits[w0 + u] : 0;
            lb[u] = G.lab[min((w0 + u) * 64 + lane, G.n - 1)];
        }
        // the run holding tuple w0 * 64 when it began before: its start and whether it is joined
        uint64_t cs = 0;
        bool cjoin = false;
        if (!(st[0] & 1) && w0 > 0) {
            uint64_t q = w0 - 1, m = st_bits[q];
            while (!m && q > 0) m = st_bits[--q];
            if (m) {
                const int bit = 63 - __builtin_clzll(m);
                cs = q * 64 + bit;
                cjoin = (rs_bits[q] >> bit) & 1;
            }
        }
        uint32_t b[kUW];
        bool live[kUW];
        int sbit[kUW];
        uint64_t ucs[kUW];  // start of the run holding the word's first tuple, for a run that began before it
#pragma unroll
        for (int u = 0; u < kUW; u++) {
            if (u > 0 && !(st[u] & 1)) {  // carried from the words before in this step, or from before the step
                bool found = false;
#pragma unroll
                for (int v = u - 1; v >= 0; v--) {
                    if (!found && st[v]) {
                        const int bit = 63 - __builtin_clzll(st[v]);
                        ucs[u] = (w0 + v) * 64 + bit;
                        found = true;
                    }
                }
                if (!found) ucs[u] = cs;
            } else {
                ucs[u] = cs;
            }
            const uint64_t sm = st[u] & upto;
            sbit[u] = sm ? 63 - __builtin_clzll(sm) : -1;  // start of the lane's run in this word
            bool join;
            if (sbit[u] >= 0) {
                join = (rs[u] >> sbit[u]) & 1;
            } else if (u == 0 || ucs[u] < w0 * 64) {
                join = cjoin;
            } else {
                const uint64_t v = ucs[u] / 64 - w0;
                join = (rs[v] >> (ucs[u] % 64)) & 1;
            }
            const uint64_t t = (w0 + u) * 64 + lane;
            b[u] = t < G.n && join ? lb[u] : 0;
            live[u] = t < G.n && join;
        }
        // the BAD words of every lane's labels loaded together (b = 0 when not
        // joined: word 0, in bounds), not one after another behind live[u]
        uint64_t bw[kUW];
#pragma unroll
        for (int u = 0; u < kUW; u++) bw[u] = badbits[b[u] >> 6];
#pragma unroll
        for (int u = 0; u < kUW; u++) live[u] = live[u] && !((bw[u] >> (b[u] & 63)) & 1);
#pragma unroll
        for (int u = 0; u < kUW; u++) {
            uint64_t cand = __ballot(live[u]) & lt;
            if (sbit[u] > 0) cand &= ~((1ull << sbit[u]) - 1);  // lanes of the same run
            const int p = cand ? 63 - __builtin_clzll(cand) : (int)lane;
            const uint32_t a = __shfl(b[u], p, kWave);
            if (!live[u]) continue;
            if (cand) {
                if (a != b[u]) uf_union(G.parent, a, b[u]);
            } else if (sbit[u] < 0) {  // first live tuple of an earlier run in this word: the run's live tuple before it
                for (uint64_t x = (w0 + u) * 64; x > ucs[u];) {
                    const uint32_t c = G.lab[--x];
                    if (!is_bad(badbits, c)) {
                        if (c != b[u]) uf_union(G.parent, c, b[u]);
                        break;
                    }
                }
            }
        }
    }
}

__device__ __forceinline__ bool zeroed_at(const GParams &G, uint64_t t) { return (G.zbits[t >> 5] >> (t & 31)) & 1; }

// The tuples are sorted by k-mer alone (stable, so each run is in generation
// order); the reference's order also sorts a run by adj.  Where that order
// matters, the first tuple of a run in the reference's order is the one with
// the smallest adj, the earliest on ties.

// First surviving tuple of each segment in the reference's order: within the
// first run holding a survivor (label not BAD, k-mer neither 0 nor zeroed),
// the survivor with the smallest adj.
__global__ void first_alive_kernel(GParams G, uint64_t *__restrict__ first)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= G.T) return;
    uint64_t i = G.seg[t];
    const uint64_t e = G.seg[t + 1];
    first[t] = ~0ull;
    while (i < e) {
        const uint64_t k = G.keys[i];
        const bool dead = k == 0 || zeroed_at(G, i);
        uint64_t j = i, best = ~0ull;
        uint32_t ba = 256;
        for (; j < e && G.keys[j] == k; j++)
            if (!dead && G.adj[j] < ba && !G.bad[G.lab[j]]) {
                ba = G.adj[j];
                best = j;
            }
        if (best != ~0ull) {
            first[t] = best;
            return;
        }
        i = j;
    }
}

__global__ void quirk_union_kernel(uint32_t *parent, uint32_t a, uint32_t b) { uf_union(parent, a, b); }

// Labels to roots in two kernels.  halve_kernel walks every read to its root
// with path halving (deep components get short paths, the walks share work);
// compress_kernel then walks the short paths without writing on the way and
// stores each read's root.  Halving writes in the same kernel as the final
// stores could put a stale ancestor back over a read's stored root.
__global__ void halve_kernel(uint32_t *parent, const uint8_t *__restrict__ bad, uint64_t N)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || (bad && bad[i])) return;
    (void)uf_find(parent, (uint32_t)i);
}

__global__ void compress_kernel(uint32_t *parent, const uint8_t *__restrict__ bad, uint64_t N)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || (bad && bad[i])) return;
    uint32_t x = (uint32_t)i;
    for (;;) {
        const uint32_t p = parent[x];
        if (p == x) break;
        x = p;
    }
    parent[i] = x;
}

// sizes[r] += reads x with parent[x] = r.  A component holding much of the
// input (the 3-minimizer graph before its pair step, a 100 M-read scan's
// largest component) would serialise millions of device atomics on one
// address, so a workgroup first counts its 256 reads' roots in a small LDS
// hash table (four probes; a root that finds no slot adds singly) and then
// adds each root it saw once.
constexpr uint32_t kSizeSlots = 512;
__global__ __launch_bounds__(256) void sizes_kernel(const uint32_t *__restrict__ parent, uint32_t *__restrict__ sizes,
                                                    uint64_t N)
{
    __shared__ uint32_t s_key[kSizeSlots], s_cnt[kSizeSlots];
    for (uint32_t j = threadIdx.x; j < kSizeSlots; j += blockDim.x) {
        s_key[j] = ~0u;
        s_cnt[j] = 0;
    }
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) {
        const uint32_t v = parent[i];
        const uint32_t h = (v * 2654435761u) >> 23;  // 9 bits
        bool done = false;
        for (uint32_t p = 0; p < 4 && !done; p++) {
            const uint32_t slot = (h + p) & (kSizeSlots - 1);
            const uint32_t old = atomicCAS(&s_key[slot], ~0u, v);
            if (old == ~0u || old == v) {
                atomicAdd(&s_cnt[slot], 1u);
                done = true;
            }
        }
        if (!done) atomicAdd(&sizes[v], 1u);
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < kSizeSlots; j += blockDim.x)
        if (s_key[j] != ~0u) atomicAdd(&sizes[s_key[j]], s_cnt[j]);
}

__global__ __launch_bounds__(256) void root_weight_kernel(const uint32_t *__restrict__ parent,
                                                          const uint32_t *__restrict__ sizes, uint32_t *__restrict__ w,
                                                          uint64_t N, unsigned long long *__restrict__ st)
{
    // grid-stride; then workgroup totals: number of roots and the largest size, one atomic each
    uint32_t nroot = 0, mx = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        const bool root = parent[i] == (uint32_t)i;
        const uint32_t s = root ? sizes[i] : 0;
        w[i] = s;
        nroot += root;
        mx = max(mx, s);
    }
    __shared__ uint32_t s_n[256 / kWave], s_m[256 / kWave];
    for (int d = kWave / 2; d > 0; d >>= 1) {
        nroot += __shfl_down(nroot, d, kWave);
        mx = max(mx, (uint32_t)__shfl_down(mx, d, kWave));
    }
    if (__lane_id() == 0) {
        s_n[threadIdx.x / kWave] = nroot;
        s_m[threadIdx.x / kWave] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a = 0, b = 0;
        for (int k = 0; k < (int)(blockDim.x / kWave); k++) {
            a += s_n[k];
            b = max(b, s_m[k]);
        }
        if (a) atomicAdd(&st[0], (unsigned long long)a);
        if (b) atomicMax(&st[1], (unsigned long long)b);
    }
}

__global__ void plan_kernel(const uint64_t *__restrict__ excl, uint64_t N, uint64_t *__restrict__ W)
{
    const int p = threadIdx.x;
    if (p > MP_NPART) return;
    if (p == 0) { W[0] = 0; return; }
    if (p == MP_NPART) { W[p] = N; return; }
    const uint64_t target = (uint64_t)p * (N / MP_NPART);  // metapart.c:4036-4053
    uint64_t lo = 0, hi = N;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (excl[mid] < target) lo = mid + 1; else hi = mid;
    }
    W[p] = lo;
}

// ---- pair step (unionFindPeReads, metapart.c:2810-2860) on compressed labels:
// parent[x] is the root of x's component, sizes[r] its read count.
__device__ __forceinline__ bool pair_candidate(const uint32_t *parent, const uint32_t *sizes, uint64_t j, uint32_t &a,
                                               uint32_t &b)
{
    a = parent[2 * j];
    b = parent[2 * j + 1];
    return a != b && (sizes[a] <= MP_PAIR_GATE || sizes[b] <= MP_PAIR_GATE);
}

// components of the candidate-pair graph, rooted at their largest id
__global__ void pair_graph_kernel(const uint32_t *__restrict__ parent, const uint32_t *__restrict__ sizes,
                                  uint32_t *gparent, uint64_t P)
{
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t a, b;
    if (j < P && pair_candidate(parent, sizes, j, a, b)) uf_union(gparent, a, b);
}

// 1 for a candidate pair inside a candidate-graph component of > 1000 reads
__global__ void pair_classify_kernel(const uint32_t *__restrict__ parent, const uint32_t *__restrict__ sizes,
                                     const uint32_t *__restrict__ gparent, const uint32_t *__restrict__ gsize,
                                     uint32_t *__restrict__ ordered, uint64_t P)
{
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P) return;
    uint32_t a, b;
    ordered[j] = pair_candidate(parent, sizes, j, a, b) && gsize[gparent[a]] > MP_PAIR_GATE;
}

// The ordered pass works on compact ids: the roots of the candidate-graph
// components of > 1000 reads, numbered in id order, so "the larger id becomes
// the root" is the same comparison on compact ids.
__global__ void big_root_kernel(const uint32_t *__restrict__ parent, const uint32_t *__restrict__ gparent,
                                const uint32_t *__restrict__ gsize, uint32_t *__restrict__ flag, uint64_t N)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < N) flag[x] = parent[x] == (uint32_t)x && gsize[gparent[x]] > MP_PAIR_GATE;
}

// roots of the candidate-graph components of > 1000 reads (numbered by a scan)
__global__ void big_groot_kernel(const uint32_t *__restrict__ gparent, const uint32_t *__restrict__ gsize,
                                 uint32_t *__restrict__ flag, uint64_t N)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < N) flag[x] = gparent[x] == (uint32_t)x && gsize[x] > MP_PAIR_GATE;
}

// ordered pairs, in pair order, as (compact id of read 2j's root, of read 2j+1's root),
// with the number of their candidate-graph component and their position
__global__ void pair_list_kernel(const uint32_t *__restrict__ ordered, const uint64_t *__restrict__ off, uint64_t P,
                                 const uint32_t *__restrict__ parent, const uint64_t *__restrict__ cid,
                                 const uint32_t *__restrict__ gparent, const uint64_t *__restrict__ gid,
                                 uint64_t *__restrict__ pairs, uint64_t *__restrict__ comp, uint32_t *__restrict__ pos)
{
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < P && ordered[j]) {
        const uint64_t q = off[j];
        const uint32_t a = parent[2 * j];
        pairs[q] = (uint64_t)(uint32_t)cid[a] | ((uint64_t)(uint32_t)cid[parent[2 * j + 1]] << 32);
        comp[q] = gid[gparent[a]];
        pos[q] = (uint32_t)q;
    }
}

// the ordered pairs grouped by component (pair order kept inside one), and
// each component's range [cf, cl) of them
__global__ void pair_group_kernel(const uint64_t *__restrict__ comp, const uint32_t *__restrict__ pos,
                                  const uint64_t *__restrict__ pairs, uint64_t n, uint64_t *__restrict__ out,
                                  uint32_t *__restrict__ cf, uint32_t *__restrict__ cl)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = pairs[pos[i]];
    const uint64_t g = comp[i];
    if (i == 0 || comp[i - 1] != g) cf[g] = (uint32_t)i;
    if (i + 1 == n || comp[i + 1] != g) cl[g] = (uint32_t)(i + 1);
}

__global__ void compact_roots_kernel(const uint32_t *__restrict__ flag, const uint64_t *__restrict__ cid,
                                     const uint32_t *__restrict__ sizes, uint32_t *__restrict__ csize,
                                     uint32_t *__restrict__ corig, uint64_t N)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < N && flag[x]) {
        csize[cid[x]] = sizes[x];
        corig[cid[x]] = (uint32_t)x;
    }
}

// Reads of candidate-graph components of <= 1000 reads take the component's
// largest id; every other initial root of such a component is one union.
__global__ void pair_apply_kernel(uint32_t *__restrict__ parent, const uint32_t *__restrict__ gparent,
                                  const uint32_t *__restrict__ gsize, uint64_t N, unsigned long long *__restrict__ unions)
{
    uint32_t cnt = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        const uint32_t g = gparent[i];
        if (gsize[g] <= MP_PAIR_GATE) {
            cnt += parent[i] == (uint32_t)i && g != (uint32_t)i;
            parent[i] = g;
        }
    }
    block_add(cnt, unions);
}

// Reads of the larger candidate-graph components take the roots the ordered
// host pass gave their initial roots (cfinal: compact id -> final compact root).
__global__ void pair_merge_kernel(uint32_t *__restrict__ parent, const uint32_t *__restrict__ gparent,
                                  const uint32_t *__restrict__ gsize, const uint64_t *__restrict__ cid,
                                  const uint32_t *__restrict__ cfinal, const uint32_t *__restrict__ corig, uint64_t N)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N && gsize[gparent[i]] > MP_PAIR_GATE) parent[i] = corig[cfinal[cid[parent[i]]]];
}

// Per-sample partition plans of a batch: W[p] over the sample's own read ids.
__global__ void sample_plan_kernel(const uint64_t *__restrict__ excl, const uint64_t *__restrict__ first,
                                   uint64_t *__restrict__ W)
{
    const int s = blockIdx.x, p = threadIdx.x;
    if (p > MP_NPART) return;
    const uint64_t off = first[s], n = first[s + 1] - off, base = excl[off];
    uint64_t *w = W + (size_t)s * (MP_NPART + 1);
    if (p == 0) { w[0] = 0; return; }
    if (p == MP_NPART) { w[p] = n; return; }
    const uint64_t target = (uint64_t)p * (n / MP_NPART);
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (excl[off + mid] - base < target) lo = mid + 1; else hi = mid;
    }
    w[p] = lo;
}

// Per-sample root counts and largest components of a batch.
__global__ __launch_bounds__(256) void sample_stats_kernel(const uint32_t *__restrict__ parent,
                                                           const uint32_t *__restrict__ sizes,
                                                           const uint64_t *__restrict__ first, int ns, uint64_t N,
                                                           unsigned long long *__restrict__ out)
{
    __shared__ uint32_t s_n[4], s_m[4];
    const uint64_t b0 = (uint64_t)blockIdx.x * blockDim.x;
    const uint64_t i = b0 + threadIdx.x;
    const uint64_t bl = min(b0 + blockDim.x, N) - 1;
    int s = 0, hi = ns - 1;
    const uint64_t x = i < N ? i : bl;
    while (s < hi) {
        const int mid = (s + hi + 1) >> 1;
        if (first[mid] <= x) s = mid; else hi = mid - 1;
    }
    const bool root = i < N && parent[i] == (uint32_t)i;
    const uint32_t sz = root ? sizes[i] : 0;
    // one sample in the whole workgroup (the common case): reduce, one atomic each
    const bool uniform = __syncthreads_and(first[s] <= b0 && bl < first[s + 1]);
    if (uniform) {
        const uint64_t m = __ballot(root);
        uint32_t mx = sz;
        for (int d = kWave / 2; d > 0; d >>= 1) mx = max(mx, (uint32_t)__shfl_down(mx, d, kWave));
        if (__lane_id() == 0) {
            s_n[threadIdx.x / kWave] = (uint32_t)__popcll(m);
            s_m[threadIdx.x / kWave] = mx;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t a = 0, b = 0;
            for (int w = 0; w < 4; w++) {
                a += s_n[w];
                b = max(b, s_m[w]);
            }
            if (a) atomicAdd(&out[2 * s], (unsigned long long)a);
            if (b) atomicMax(&out[2 * s + 1], (unsigned long long)b);
        }
    } else if (root) {
        atomicAdd(&out[2 * s], 1ull);
        atomicMax(&out[2 * s + 1], (unsigned long long)sz);
    }
}

static inline unsigned nblk(uint64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// every read (except BAD ones, when bad is given) to its root
static int compress_labels(mp_ctx *ctx, uint32_t *parent, const uint8_t *bad, uint64_t N)
{
    if (N == 0) return 0;
    hipLaunchKernelGGL(halve_kernel, dim3(nblk(N)), dim3(256), 0, ctx->stream, parent, bad, N);
    hipLaunchKernelGGL(compress_kernel, dim3(nblk(N)), dim3(256), 0, ctx->stream, parent, bad, N);
    MP_HIP(hipGetLastError());
    return 0;
}

// Per segment: first surviving tuple, its k-mer and label, and the last k-mer
// (0 if zeroed), first label and that label's BAD flag of the segment's quirk
// group, all in the reference's tuple order.
__global__ void quirk_info_kernel(GParams G, const uint64_t *__restrict__ first, const uint64_t *__restrict__ grp,
                                  uint64_t *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= G.T) return;
    const uint64_t f = first[t], gs = grp[2 * t], ge = grp[2 * t + 1];
    uint64_t *o = out + 5 * (uint64_t)t;
    o[0] = f;
    o[1] = f != ~0ull ? G.keys[f] : 0;
    o[2] = f != ~0ull ? G.lab[f] : 0;
    o[3] = ge > gs && !zeroed_at(G, ge - 1) ? G.keys[ge - 1] : 0;
    uint32_t L0 = 0;
    if (ge > gs) {  // tuple 0 of the group: smallest adj of its first run
        const uint64_t k = G.keys[gs];
        uint32_t ba = 256;
        for (uint64_t j = gs; j < ge && G.keys[j] == k; j++)
            if (G.adj[j] < ba) {
                ba = G.adj[j];
                L0 = G.lab[j];
            }
    }
    o[4] = (uint64_t)L0 | ((uint64_t)(ge > gs && G.bad[L0]) << 32) | ((uint64_t)(ge > gs) << 33);
}

// Filter, union and compress over the n sorted tuples of one scan, split into
// nseg segments at device positions d_seg[0..nseg].  grp (host, 2 per segment)
// is the tuple range whose last k-mer and first label drive the segment's
// first-run quirk: the whole scan for one sample, the sample's own tuples in
// a batch of samples.
static int filter_union_core(mp_ctx *ctx, uint64_t n, const uint64_t *d_seg, int nseg, const uint64_t *grp,
                             mp_scan_stats *st)
{
    const uint64_t N = ctx->N;
    uint64_t *keys = (uint64_t *)ctx->keys[ctx->cur].p;
    const uint32_t *lab = (const uint32_t *)ctx->lab[ctx->cur].p;
    const uint8_t *adj = (const uint8_t *)ctx->adj[ctx->cur].p;
    const uint64_t nwords = (n + 31) / 32 + 1;
    MP_TRY(ensure(ctx, ctx->zbits, nwords * sizeof(uint32_t)));
    MP_TRY(fill_u32(ctx, ctx->zbits.p, 0, nwords));
    // small device scratch: first-alive, quirk groups and info, stats
    uint64_t *d_small = nullptr;
    const size_t small_words = 8 * (size_t)nseg + 16;
    MP_HIP(hipMallocAsync((void **)&d_small, small_words * sizeof(uint64_t), ctx->stream));
    uint64_t *d_first = d_small, *d_grp = d_small + nseg, *d_info = d_small + 3 * nseg;
    unsigned long long *d_stats = (unsigned long long *)(d_small + 8 * nseg);
    MP_TRY(fill_u32(ctx, d_stats, 0, 32));
    MP_HIP(hipMemcpyAsync(d_grp, grp, sizeof(uint64_t) * 2 * nseg, hipMemcpyHostToDevice, ctx->stream));
    GParams G{};
    G.keys = keys;
    G.adj = adj;
    G.lab = lab;
    G.n = n;
    G.seg = d_seg;
    G.T = nseg;
    G.bad = (uint8_t *)ctx->bad.p;
    G.filtered = (uint8_t *)ctx->filtered.p;
    G.zbits = (uint32_t *)ctx->zbits.p;
    G.parent = (uint32_t *)ctx->parent.p;
    G.stats = d_stats;
    hipEventRecord(ctx->ev[2], ctx->stream);
    const uint64_t rwords = (n + 63) / 64 + 1;
    MP_TRY(ensure(ctx, ctx->runbits, 2 * rwords * sizeof(uint64_t)));
    uint64_t *rs_bits = (uint64_t *)ctx->runbits.p, *st_bits = rs_bits + rwords;
    if (!grid_fits((n + kFTile - 1) / kFTile, kFLanes))
        return ctx->fail(MP_EINVAL, "filter over %llu tuples exceeds one dispatch: use more scans", (unsigned long long)n);
    if (n) {
        hipLaunchKernelGGL(filter_tile_kernel, dim3((unsigned)((n + kFTile - 1) / kFTile)), dim3(kFLanes), 0,
                           ctx->stream, G, rs_bits, st_bits);
        MP_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(refilter_kernel, dim3(stride_grid(N)), dim3(256), 0, ctx->stream, G.bad,
                       (const uint8_t *)G.filtered, G.parent, N, d_stats + 2);
    MP_HIP(hipGetLastError());
    hipEventRecord(ctx->ev[3], ctx->stream);
    // first-run quirk (metapart.c:2539-2616): each segment starts with its
    // group's last k-mer as the "current" k-mer and the group's first label as
    // the centre, so a segment whose first surviving run has that k-mer joins
    // its labels to that centre
    int quirks = 0;
    if (n) {
        hipLaunchKernelGGL(first_alive_kernel, dim3(nblk(nseg)), dim3(256), 0, ctx->stream, G, d_first);
        hipLaunchKernelGGL(quirk_info_kernel, dim3(nblk(nseg)), dim3(256), 0, ctx->stream, G,
                           (const uint64_t *)d_first, (const uint64_t *)d_grp, d_info);
        MP_HIP(hipGetLastError());
        std::vector<uint64_t> info(5 * (size_t)nseg);
        MP_HIP(hipMemcpyAsync(info.data(), d_info, sizeof(uint64_t) * 5 * nseg, hipMemcpyDeviceToHost, ctx->stream));
        MP_HIP(hipStreamSynchronize(ctx->stream));
        for (int t = 0; t < nseg; t++) {
            const uint64_t *o = &info[5 * (size_t)t];
            if (o[0] == ~0ull || !((o[4] >> 33) & 1) || o[1] != o[3]) continue;
            if ((o[4] >> 32) & 1) {
                MP_HIP(hipFreeAsync(d_small, ctx->stream));
                return ctx->fail(MP_EPLAN, "first-run union quirk with a filtered tuple-0 label (segment %d)", t);
            }
            quirks++;
            const uint32_t L0 = (uint32_t)o[4], lf = (uint32_t)o[2];
            if (lf != L0) {
                hipLaunchKernelGGL(quirk_union_kernel, dim3(1), dim3(1), 0, ctx->stream, G.parent, L0, lf);
                MP_HIP(hipGetLastError());
            }
        }
        MP_TRY(ensure(ctx, ctx->scratch, ((N + 63) / 64 + 1) * sizeof(uint64_t)));
        uint64_t *badbits = (uint64_t *)ctx->scratch.p;
        hipLaunchKernelGGL(bad_bits_kernel, dim3(nblk(N)), dim3(256), 0, ctx->stream, (const uint8_t *)G.bad, N,
                           badbits);
        MP_HIP(hipGetLastError());
        hipLaunchKernelGGL(union_kernel, dim3(wave_grid(((n + 63) / 64 + kUW - 1) / kUW)), dim3(256), 0, ctx->stream,
                           G, (const uint64_t *)rs_bits, (const uint64_t *)st_bits, (const uint64_t *)badbits);
        MP_HIP(hipGetLastError());
    }
    MP_TRY(compress_labels(ctx, G.parent, (const uint8_t *)G.bad, N));
    MP_HIP(hipGetLastError());
    hipEventRecord(ctx->ev[4], ctx->stream);
    unsigned long long hs[4] = {0, 0, 0, 0};
    MP_HIP(hipMemcpyAsync(hs, d_stats, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipFreeAsync(d_small, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    if (st) {
        st->evaluated_runs = hs[0];
        st->zeroed_runs = hs[1];
        st->bad_labels = hs[2];
        st->quirk_segments = quirks;
        float a = 0, b = 0;
        hipEventElapsedTime(&a, ctx->ev[2], ctx->ev[3]);
        hipEventElapsedTime(&b, ctx->ev[3], ctx->ev[4]);
        st->ms_filter = a;
        st->ms_union = b;
    }
    return 0;
}

int filter_union_scan(mp_ctx *ctx, uint64_t n, const uint64_t *seg_bounds, int T, mp_scan_stats *st)
{
    const int bshift = 2 * (ctx->cfg.k - 10);
    uint64_t *d_small = nullptr;
    MP_HIP(hipMallocAsync((void **)&d_small, 2 * (size_t)(T + 1) * sizeof(uint64_t), ctx->stream));
    uint64_t *d_bounds = d_small, *d_seg = d_small + (T + 1);
    MP_HIP(hipMemcpyAsync(d_bounds, seg_bounds, sizeof(uint64_t) * (T + 1), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(seg_pos_kernel, dim3(nblk(T + 1)), dim3(256), 0, ctx->stream,
                       (const uint64_t *)ctx->keys[ctx->cur].p, n, bshift, (const uint64_t *)d_bounds, T, d_seg);
    MP_HIP(hipGetLastError());
    std::vector<uint64_t> segpos(T + 1);
    MP_HIP(hipMemcpyAsync(segpos.data(), d_seg, sizeof(uint64_t) * (T + 1), hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    if (segpos[0] != 0 || segpos[T] != n) {
        MP_HIP(hipFreeAsync(d_small, ctx->stream));
        return ctx->fail(MP_EPLAN, "tuples outside the scan's thread ranges");
    }
    std::vector<uint64_t> grp(2 * (size_t)T);
    for (int t = 0; t < T; t++) {
        grp[2 * t] = 0;
        grp[2 * t + 1] = n;
    }
    const int rc = filter_union_core(ctx, n, d_seg, T, grp.data(), st);
    MP_HIP(hipFreeAsync(d_small, ctx->stream));
    return rc;
}

// ---------------------------------------------------------------- two sweeps
// A scan whose tuples do not fit the device budget, over bucket parts that do
// (cuts: part boundaries in bucket space, from lo to hi).  The filter's BAD set
// must be complete before any union of the scan, so the parts go by twice:
//   sweep 1: each part's tuples are made and sorted, and filtered (filtered
//            marks, evaluated and zeroed runs); the scan's first label (tuple
//            0's, from the first part) and last k-mer (from the last part, 0
//            if zeroed) are kept for the first-run quirk; then the refilter
//            turns the marks into the BAD set;
//   sweep 2: each part's tuples are made, sorted and filtered again (for the
//            zeroed runs and the run bitmaps; the marks repeat sweep 1's), the
//            first surviving run of every segment not yet met takes the quirk,
//            and the part's runs are joined.
// Runs never cross a bucket, so never a part; a segment cut by a part end
// keeps its last run evaluated there (open_last).  Joins in any order give the
// same components, so the result is the one-pass scan's.
int two_sweep_scan(mp_ctx *ctx, const std::vector<uint64_t> &cuts, const uint64_t *seg_bounds, int T,
                   const std::function<int(uint64_t, uint64_t, uint64_t *)> &make_part, mp_scan_stats *st)
{
    const uint64_t N = ctx->N;
    const int bshift = 2 * (ctx->cfg.k - 10);
    const int P = (int)cuts.size() - 1;
    uint64_t *d_small = nullptr;
    const size_t small_words = 9 * (size_t)T + 2 * (size_t)(T + 1) + 32;
    MP_HIP(hipMallocAsync((void **)&d_small, small_words * sizeof(uint64_t), ctx->stream));
    uint64_t *d_first = d_small, *d_grp = d_small + T, *d_info = d_small + 3 * T, *d_bounds = d_small + 8 * T,
             *d_seg = d_bounds + (T + 1);
    unsigned long long *d_stats = (unsigned long long *)(d_seg + (T + 1));
    MP_TRY(fill_u32(ctx, d_stats, 0, 2 * 32));
    GParams G{};
    G.bad = (uint8_t *)ctx->bad.p;
    G.filtered = (uint8_t *)ctx->filtered.p;
    G.parent = (uint32_t *)ctx->parent.p;
    bool have_L0 = false, have_last = false;
    uint32_t L0 = 0;
    uint64_t lastK = 0, total = 0;
    std::vector<char> met(T, 0);
    int quirks = 0;
    float ms_filter = 0, ms_union = 0, ms = 0;
    // one part: its tuples sorted, its local segments, zeroed runs and run bitmaps
    auto prepare = [&](int p, int sweep, uint64_t *n_out, int *nloc_out, std::vector<int> &segid) -> int {
        const uint64_t plo = cuts[p], phi = cuts[p + 1];
        uint64_t n = 0;
        MP_TRY(make_part(plo, phi, &n));
        *n_out = n;
        segid.clear();
        std::vector<uint64_t> lb;
        for (int t = 0; t < T; t++) {
            if (seg_bounds[t + 1] <= plo || seg_bounds[t] >= phi) continue;
            lb.push_back(std::max(seg_bounds[t], plo));
            segid.push_back(t);
        }
        lb.push_back(std::min(seg_bounds[segid.back() + 1], phi));
        const int nloc = (int)segid.size();
        *nloc_out = nloc;
        if (n == 0) return 0;
        MP_HIP(hipMemcpyAsync(d_bounds, lb.data(), sizeof(uint64_t) * (nloc + 1), hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(seg_pos_kernel, dim3(nblk(nloc + 1)), dim3(256), 0, ctx->stream,
                           (const uint64_t *)ctx->keys[ctx->cur].p, n, bshift, (const uint64_t *)d_bounds, nloc, d_seg);
        MP_HIP(hipGetLastError());
        G.keys = (uint64_t *)ctx->keys[ctx->cur].p;
        G.lab = (const uint32_t *)ctx->lab[ctx->cur].p;
        G.adj = (const uint8_t *)ctx->adj[ctx->cur].p;
        G.n = n;
        G.seg = d_seg;
        G.T = nloc;
        G.open_last = seg_bounds[segid.back() + 1] > phi;
        G.stats = d_stats + (sweep == 1 ? 0 : 16);  // the second sweep's counts repeat the first's
        const uint64_t nwords = (n + 31) / 32 + 1, rwords = (n + 63) / 64 + 1;
        MP_TRY(ensure(ctx, ctx->zbits, nwords * sizeof(uint32_t)));
        MP_TRY(ensure(ctx, ctx->runbits, 2 * rwords * sizeof(uint64_t)));
        G.zbits = (uint32_t *)ctx->zbits.p;
        MP_TRY(fill_u32(ctx, ctx->zbits.p, 0, nwords));
        if (!grid_fits((n + kFTile - 1) / kFTile, kFLanes))
            return ctx->fail(MP_EINVAL, "filter over %llu tuples exceeds one dispatch", (unsigned long long)n);
        hipEventRecord(ctx->ev[2], ctx->stream);
        uint64_t *rs_bits = (uint64_t *)ctx->runbits.p;
        hipLaunchKernelGGL(filter_tile_kernel, dim3((unsigned)((n + kFTile - 1) / kFTile)), dim3(kFLanes), 0,
                           ctx->stream, G, rs_bits, rs_bits + rwords);
        MP_HIP(hipGetLastError());
        hipEventRecord(ctx->ev[3], ctx->stream);
        // first survivor of every local segment and its k-mer and label; the part's
        // first label and last k-mer (0 if zeroed)
        std::vector<uint64_t> grp(2 * (size_t)nloc);
        for (int l = 0; l < nloc; l++) {
            grp[2 * l] = 0;
            grp[2 * l + 1] = n;
        }
        MP_HIP(hipMemcpyAsync(d_grp, grp.data(), sizeof(uint64_t) * 2 * nloc, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(first_alive_kernel, dim3(nblk(nloc)), dim3(256), 0, ctx->stream, G, d_first);
        hipLaunchKernelGGL(quirk_info_kernel, dim3(nblk(nloc)), dim3(256), 0, ctx->stream, G,
                           (const uint64_t *)d_first, (const uint64_t *)d_grp, d_info);
        MP_HIP(hipGetLastError());
        MP_HIP(hipStreamSynchronize(ctx->stream));
        hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]);
        ms_filter += ms;
        return 0;
    };
    std::vector<int> segid;
    for (int p = 0; p < P; p++) {  // sweep 1
        uint64_t n = 0;
        int nloc = 0;
        MP_TRY(prepare(p, 1, &n, &nloc, segid));
        total += n;
        if (!n) continue;
        std::vector<uint64_t> info(5 * (size_t)nloc);
        MP_HIP(hipMemcpy(info.data(), d_info, sizeof(uint64_t) * 5 * nloc, hipMemcpyDeviceToHost));
        if (!have_L0) {
            L0 = (uint32_t)info[4];
            have_L0 = true;
        }
        lastK = info[3];
        have_last = true;
    }
    hipEventRecord(ctx->ev[2], ctx->stream);
    hipLaunchKernelGGL(refilter_kernel, dim3(stride_grid(N)), dim3(256), 0, ctx->stream, G.bad,
                       (const uint8_t *)G.filtered, G.parent, N, d_stats + 2);
    MP_HIP(hipGetLastError());
    hipEventRecord(ctx->ev[3], ctx->stream);
    MP_HIP(hipStreamSynchronize(ctx->stream));
    hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]);
    ms_filter += ms;
    uint8_t bad_L0 = 0;
    if (have_L0) MP_HIP(hipMemcpy(&bad_L0, G.bad + L0, 1, hipMemcpyDeviceToHost));
    for (int p = 0; p < P; p++) {  // sweep 2
        uint64_t n = 0;
        int nloc = 0;
        MP_TRY(prepare(p, 2, &n, &nloc, segid));
        if (!n) continue;
        std::vector<uint64_t> info(5 * (size_t)nloc);
        MP_HIP(hipMemcpy(info.data(), d_info, sizeof(uint64_t) * 5 * nloc, hipMemcpyDeviceToHost));
        hipEventRecord(ctx->ev[4], ctx->stream);
        for (int l = 0; l < nloc; l++) {  // metapart.c:2539-2616, as in the one-pass scan
            const uint64_t *o = &info[5 * (size_t)l];
            const int t = segid[l];
            if (met[t] || o[0] == ~0ull) continue;
            met[t] = 1;
            if (!have_last || o[1] != lastK) continue;
            if (bad_L0) {
                MP_HIP(hipFreeAsync(d_small, ctx->stream));
                return ctx->fail(MP_EPLAN, "first-run union quirk with a filtered tuple-0 label (segment %d)", t);
            }
            quirks++;
            const uint32_t lf = (uint32_t)o[2];
            if (lf != L0) {
                hipLaunchKernelGGL(quirk_union_kernel, dim3(1), dim3(1), 0, ctx->stream, G.parent, L0, lf);
                MP_HIP(hipGetLastError());
            }
        }
        // the BAD set as bits (the sort of this part used the scratch array)
        MP_TRY(ensure(ctx, ctx->scratch, ((N + 63) / 64 + 1) * sizeof(uint64_t)));
        uint64_t *badbits = (uint64_t *)ctx->scratch.p;
        hipLaunchKernelGGL(bad_bits_kernel, dim3(nblk(N)), dim3(256), 0, ctx->stream, (const uint8_t *)G.bad, N,
                           badbits);
        const uint64_t rwords = (n + 63) / 64 + 1;
        const uint64_t *rs_bits = (const uint64_t *)ctx->runbits.p;
        hipLaunchKernelGGL(union_kernel, dim3(wave_grid(((n + 63) / 64 + kUW - 1) / kUW)), dim3(256), 0,
                           ctx->stream, G, rs_bits, rs_bits + rwords, (const uint64_t *)badbits);
        MP_HIP(hipGetLastError());
        hipEventRecord(ctx->ev[5], ctx->stream);
        MP_HIP(hipStreamSynchronize(ctx->stream));
        hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]);
        ms_union += ms;
    }
    MP_TRY(compress_labels(ctx, G.parent, (const uint8_t *)G.bad, N));
    unsigned long long hs[4] = {0, 0, 0, 0};
    MP_HIP(hipMemcpyAsync(hs, d_stats, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipFreeAsync(d_small, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    if (st) {
        st->tuples = total;
        st->evaluated_runs = hs[0];
        st->zeroed_runs = hs[1];
        st->bad_labels = hs[2];
        st->quirk_segments = quirks;
        st->ms_filter = ms_filter;
        st->ms_union = ms_union;
        st->parts = P;
    }
    return 0;
}

int filter_union_groups(mp_ctx *ctx, uint64_t n, const uint64_t *seg_pos, int nseg, const uint64_t *grp,
                        mp_scan_stats *st)
{
    uint64_t *d_seg = nullptr;
    MP_HIP(hipMallocAsync((void **)&d_seg, (size_t)(nseg + 1) * sizeof(uint64_t), ctx->stream));
    MP_HIP(hipMemcpyAsync(d_seg, seg_pos, sizeof(uint64_t) * (nseg + 1), hipMemcpyHostToDevice, ctx->stream));
    const int rc = filter_union_core(ctx, n, d_seg, nseg, grp, st);
    MP_HIP(hipFreeAsync(d_seg, ctx->stream));
    return rc;
}

// ---------------------------------------------------------------- finish
// Sequential pair step (unionFindPeReads, metapart.c:2801-2867) on the host,
// for the pairs whose outcome depends on the order.
// pairs between the prefetch of a pair's nodes and their use, and between the
// prefetch of those nodes' parents (read from the now cached nodes) and their use
constexpr uint64_t kPairAhead = 16, kParentAhead = 8;

struct PairNode {  // a compact root: its parent, and its component's size while it is a root
    uint32_t p, s;
};

static uint32_t hfind(PairNode *nd, uint32_t x)
{
    while (nd[x].p != x) {
        const uint32_t p = nd[x].p, gp = nd[p].p;
        nd[x].p = gp;
        x = gp;
    }
    return x;
}

int finish_components(mp_ctx *ctx, uint64_t *W, mp_finish_stats *st)
{
    const uint64_t N = ctx->N;
    uint32_t *parent = (uint32_t *)ctx->parent.p;
    uint32_t *sizes = (uint32_t *)ctx->sizes.p;
    hipEventRecord(ctx->ev[5], ctx->stream);
    // sizes after the last scan (metapart.c:4010-4021)
    MP_TRY(compress_labels(ctx, parent, (const uint8_t *)ctx->bad.p, N));
    MP_TRY(fill_u32(ctx, sizes, 0, N));
    hipLaunchKernelGGL(sizes_kernel, dim3(nblk(N)), dim3(256), 0, ctx->stream, (const uint32_t *)parent, sizes, N);
    MP_HIP(hipGetLastError());
    // Pair step (unionFindPeReads): pair (2j, 2j+1) joins its two components
    // unless both hold more than 1000 reads at that moment.  Sizes only grow,
    // so a pair whose components are both > 1000 now never joins, and the
    // other pairs ("candidates") form a graph G over the current components.
    // A component of G holding <= 1000 reads in total can never contain a
    // component > 1000, so every candidate inside it joins whatever the order:
    // it ends as one component rooted at its largest id.  Only candidates
    // inside larger components of G depend on the order; those run the exact
    // sequential loop on the host, on their own (they share no reads with the
    // rest).
    const uint64_t P = N / 2;
    uint64_t *d_small = nullptr;
    MP_HIP(hipMallocAsync((void **)&d_small, 64 * sizeof(uint64_t), ctx->stream));
    MP_TRY(fill_u32(ctx, d_small, 0, 128));
    unsigned long long *d_unions = (unsigned long long *)(d_small + 40);
    MP_TRY(ensure(ctx, ctx->gparent, N * sizeof(uint32_t)));
    MP_TRY(ensure(ctx, ctx->gsize, N * sizeof(uint32_t)));
    MP_TRY(ensure(ctx, ctx->wg_cnt, (P + 1) * sizeof(uint32_t)));
    MP_TRY(ensure(ctx, ctx->wg_off, (P + 2) * sizeof(uint64_t)));
    uint32_t *gparent = (uint32_t *)ctx->gparent.p, *gsize = (uint32_t *)ctx->gsize.p;
    uint32_t *ordered = (uint32_t *)ctx->wg_cnt.p;
    uint64_t n_ordered = 0;
    // host side, in the ctx's pinned buffer: the ordered pairs, the compact
    // roots' sizes, the components' ranges of pairs (first [0, Kg), ends
    // [Kg, 2 Kg)) and the compact roots' final roots
    uint64_t *hpairs = nullptr;
    uint32_t *cs = nullptr, *cfl = nullptr, *cp = nullptr;
    uint64_t M = 0;   // roots in candidate-graph components of > 1000 reads
    uint64_t Kg = 0;  // those components
    uint32_t *d_flag = nullptr, *d_csize = nullptr, *d_corig = nullptr, *d_cfinal = nullptr, *d_cfl = nullptr;
    uint64_t *d_cid = nullptr, *d_pairs = nullptr;
    if (P) {
        MP_HIP(hipMemcpyAsync(gparent, parent, N * 4, hipMemcpyDeviceToDevice, ctx->stream));
        hipLaunchKernelGGL(pair_graph_kernel, dim3(nblk(P)), dim3(256), 0, ctx->stream, (const uint32_t *)parent,
                           (const uint32_t *)sizes, gparent, P);
        MP_TRY(compress_labels(ctx, gparent, nullptr, N));
        MP_TRY(fill_u32(ctx, gsize, 0, N));
        hipLaunchKernelGGL(sizes_kernel, dim3(nblk(N)), dim3(256), 0, ctx->stream, (const uint32_t *)gparent, gsize,
                           N);
        hipLaunchKernelGGL(pair_classify_kernel, dim3(nblk(P)), dim3(256), 0, ctx->stream, (const uint32_t *)parent,
                           (const uint32_t *)sizes, (const uint32_t *)gparent, (const uint32_t *)gsize, ordered, P);
        MP_HIP(hipGetLastError());
        MP_TRY(scan_u32(ctx, ordered, (uint64_t *)ctx->wg_off.p, P));
        MP_HIP(hipMemcpyAsync(&n_ordered, (uint64_t *)ctx->wg_off.p + P, sizeof(uint64_t), hipMemcpyDeviceToHost,
                              ctx->stream));
        MP_HIP(hipStreamSynchronize(ctx->stream));
        if (n_ordered) {  // the ordered pairs on compact root ids, and those roots' sizes
            MP_TRY(ensure(ctx, ctx->pairc, N * 4 + 2 * (N + 1) * 8 + n_ordered * 8 + 3 * N * 4 + 256));
            uint8_t *pc = (uint8_t *)ctx->pairc.p;
            d_flag = (uint32_t *)pc;
            d_cid = (uint64_t *)(pc + ((N * 4 + 7) & ~7ull));
            uint64_t *d_gid = d_cid + (N + 1);
            d_pairs = d_gid + (N + 1);
            d_csize = (uint32_t *)(d_pairs + n_ordered);
            d_corig = d_csize + N;
            d_cfinal = d_corig + N;
            // pairs of different candidate-graph components share no roots: they
            // are grouped by component (a stable sort), and the components run
            // on several host threads
            for (int b = 0; b < 2; b++) {
                MP_TRY(ensure(ctx, ctx->items[b], n_ordered * sizeof(uint64_t)));
                MP_TRY(ensure(ctx, ctx->ilens[b], n_ordered * sizeof(uint32_t)));
                MP_TRY(ensure(ctx, ctx->iaux[b], n_ordered + 16));
            }
            hipLaunchKernelGGL(big_groot_kernel, dim3(nblk(N)), dim3(256), 0, ctx->stream, (const uint32_t *)gparent,
                               (const uint32_t *)gsize, d_flag, N);
            MP_HIP(hipGetLastError());
            MP_TRY(scan_u32(ctx, d_flag, d_gid, N));
            hipLaunchKernelGGL(big_root_kernel, dim3(nblk(N)), dim3(256), 0, ctx->stream, (const uint32_t *)parent,
                               (const uint32_t *)gparent, (const uint32_t *)gsize, d_flag, N);
            MP_HIP(hipGetLastError());
            MP_TRY(scan_u32(ctx, d_flag, d_cid, N));
            hipLaunchKernelGGL(pair_list_kernel, dim3(nblk(P)), dim3(256), 0, ctx->stream, (const uint32_t *)ordered,
                               (const uint64_t *)ctx->wg_off.p, P, (const uint32_t *)parent, (const uint64_t *)d_cid,
                               (const uint32_t *)gparent, (const uint64_t *)d_gid, d_pairs,
                               (uint64_t *)ctx->items[0].p, (uint32_t *)ctx->ilens[0].p);
            hipLaunchKernelGGL(compact_roots_kernel, dim3(nblk(N)), dim3(256), 0, ctx->stream,
                               (const uint32_t *)d_flag, (const uint64_t *)d_cid, (const uint32_t *)sizes, d_csize,
                               d_corig, N);
            MP_HIP(hipGetLastError());
            MP_HIP(hipMemcpyAsync(&M, d_cid + N, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
            MP_HIP(hipMemcpyAsync(&Kg, d_gid + N, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
            MP_HIP(hipStreamSynchronize(ctx->stream));
            const uint32_t gmask = Kg <= 256 ? 0x2u : Kg <= 65536 ? 0x6u : Kg <= (1u << 24) ? 0xEu : 0x1Eu;
            int cur = 0, gpasses = 0;
            MP_TRY(radix_sort_generic(ctx, ctx->items, ctx->ilens, ctx->iaux, &cur, n_ordered, gmask, &gpasses,
                                      nullptr, 0, nullptr));
            MP_HIP(hipMallocAsync((void **)&d_cfl, 2 * Kg * sizeof(uint32_t), ctx->stream));
            MP_TRY(fill_u32(ctx, d_cfl, 0, 2 * Kg));
            uint64_t *d_sorted = (uint64_t *)ctx->items[1 - cur].p;
            hipLaunchKernelGGL(pair_group_kernel, dim3(nblk(n_ordered)), dim3(256), 0, ctx->stream,
                               (const uint64_t *)ctx->items[cur].p, (const uint32_t *)ctx->ilens[cur].p,
                               (const uint64_t *)d_pairs, n_ordered, d_sorted, d_cfl, d_cfl + Kg);
            MP_HIP(hipGetLastError());
            // pinned: the ctx's own buffer when it is large enough, else the FASTQ
            // upload's two staging buffers when present (idle once the upload
            // has left) and each part fits one, else a new buffer of the ctx
            // (pinning costs about 0.25 ms per MB, once)
            const uint64_t need_a = (n_ordered * 8 + 63) & ~63ull, need_b = (2 * M + 2 * Kg) * 4 + 64;
            uint8_t *pa = nullptr, *pb = nullptr;
            if (ctx->pair_host_bytes >= need_a + need_b) {
                pa = (uint8_t *)ctx->pair_host;
                pb = pa + need_a;
            } else if (ctx->stage[0] && ctx->stage[1] && need_a <= kStageBytes && need_b <= kStageBytes) {
                for (int b = 0; b < 2; b++)
                    if (ctx->stage_used[b]) MP_HIP(hipEventSynchronize(ctx->stage_ev[b]));
                pa = (uint8_t *)ctx->stage[0];
                pb = (uint8_t *)ctx->stage[1];
            } else {
                MP_HIP(hipStreamSynchronize(ctx->stream));
                if (ctx->pair_host) MP_HIP(hipHostFree(ctx->pair_host));
                ctx->pair_host = nullptr;
                ctx->pair_host_bytes = 0;
                const uint64_t want = (need_a + need_b) + (need_a + need_b) / 4;  // room to grow
                MP_HIP(hipHostMalloc(&ctx->pair_host, want, hipHostMallocDefault));
                ctx->pair_host_bytes = want;
                pa = (uint8_t *)ctx->pair_host;
                pb = pa + need_a;
            }
            hpairs = (uint64_t *)pa;
            cs = (uint32_t *)pb;
            cfl = cs + M;
            cp = cfl + 2 * Kg;
            MP_HIP(hipMemcpyAsync(hpairs, d_sorted, n_ordered * 8, hipMemcpyDeviceToHost, ctx->stream));
            MP_HIP(hipMemcpyAsync(cs, d_csize, M * 4, hipMemcpyDeviceToHost, ctx->stream));
            MP_HIP(hipMemcpyAsync(cfl, d_cfl, 2 * Kg * 4, hipMemcpyDeviceToHost, ctx->stream));
            MP_HIP(hipFreeAsync(d_cfl, ctx->stream));
        }
        hipLaunchKernelGGL(pair_apply_kernel, dim3(stride_grid(N)), dim3(256), 0, ctx->stream, parent,
                           (const uint32_t *)gparent, (const uint32_t *)gsize, N, d_unions);
        MP_HIP(hipGetLastError());
    }
    uint64_t unions = 0;
    double host_ms = 0;
    if (n_ordered) {
        MP_HIP(hipStreamSynchronize(ctx->stream));
        const auto t0 = std::chrono::steady_clock::now();
        // parent and size side by side: a find ends on the line it reads the size from
        static_assert(sizeof(PairNode) == sizeof(uint64_t), "one node per 64-bit word");
        if (ctx->pair_nodes.size() < M) ctx->pair_nodes.resize(M);  // kept by the ctx: no page faults after the first call
        PairNode *nd = (PairNode *)ctx->pair_nodes.data();
        for (uint64_t c = 0; c < M; c++) nd[c] = {(uint32_t)c, cs[c]};
        // each component's pairs in pair order, the reference's loop; the
        // components touch disjoint nodes
        std::atomic<uint64_t> next{0};
        const unsigned nth = (unsigned)std::max<uint64_t>(
            1, std::min<uint64_t>({16, std::max(1u, std::thread::hardware_concurrency()), Kg}));
        std::vector<uint64_t> tun(nth, 0);
        auto work = [&](unsigned t) {
            for (uint64_t g; (g = next.fetch_add(1)) < Kg;) {
                const uint64_t qb = cfl[g], qe = cfl[Kg + g];
                for (uint64_t q = qb; q < qe; q++) {
                    if (q + kPairAhead < qe) {  // the nodes a later pair starts from
                        const uint64_t f = hpairs[q + kPairAhead];
                        __builtin_prefetch(&nd[(uint32_t)f]);
                        __builtin_prefetch(&nd[(uint32_t)(f >> 32)]);
                    }
                    if (q + kParentAhead < qe) {  // and their parents, the second step of its finds
                        const uint64_t f = hpairs[q + kParentAhead];
                        __builtin_prefetch(&nd[nd[(uint32_t)f].p]);
                        __builtin_prefetch(&nd[nd[(uint32_t)(f >> 32)].p]);
                    }
                    const uint32_t a = hfind(nd, (uint32_t)hpairs[q]), b = hfind(nd, (uint32_t)(hpairs[q] >> 32));
                    if (a == b) continue;
                    if (nd[a].s <= MP_PAIR_GATE || nd[b].s <= MP_PAIR_GATE) {
                        tun[t]++;
                        if (a > b) { nd[b].p = a; nd[a].s += nd[b].s; nd[b].s = 0; }
                        else { nd[a].p = b; nd[b].s += nd[a].s; nd[a].s = 0; }
                    }
                }
            }
        };
        std::vector<std::thread> pool;
        try {
            for (unsigned t = 1; t < nth; t++) pool.emplace_back(work, t);
        } catch (const std::exception &) {  // fewer threads: the others take their components
        }
        work(0);
        for (auto &th : pool) th.join();
        for (unsigned t = 0; t < nth; t++) unions += tun[t];
        host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        // the final root of every compact root, by pointer jumping on the device
        for (uint64_t c = 0; c < M; c++) cp[c] = nd[c].p;
        MP_HIP(hipMemcpyAsync(d_cfinal, cp, M * 4, hipMemcpyHostToDevice, ctx->stream));
        MP_TRY(compress_labels(ctx, d_cfinal, nullptr, M));
        hipLaunchKernelGGL(pair_merge_kernel, dim3(nblk(N)), dim3(256), 0, ctx->stream, parent,
                           (const uint32_t *)gparent, (const uint32_t *)gsize, (const uint64_t *)d_cid,
                           (const uint32_t *)d_cfinal, (const uint32_t *)d_corig, N);
        MP_HIP(hipGetLastError());
        MP_HIP(hipStreamSynchronize(ctx->stream));  // cp is read by the copy above
    }
    // every BAD read was reset to itself by the pair step
    MP_TRY(fill_u32(ctx, ctx->bad.p, 0, (N + 3) / 4));
    MP_TRY(compress_labels(ctx, parent, (const uint8_t *)ctx->bad.p, N));
    MP_TRY(fill_u32(ctx, sizes, 0, N));
    hipLaunchKernelGGL(sizes_kernel, dim3(nblk(N)), dim3(256), 0, ctx->stream, (const uint32_t *)parent, sizes, N);
    MP_HIP(hipGetLastError());
    // partition plan over root ids in ascending order
    MP_TRY(ensure(ctx, ctx->wg_off, (N + 1) * sizeof(uint64_t)));
    MP_TRY(ensure(ctx, ctx->wg_cnt, N * sizeof(uint32_t)));
    hipLaunchKernelGGL(root_weight_kernel, dim3(stride_grid(N)), dim3(256), 0, ctx->stream, (const uint32_t *)parent,
                       (const uint32_t *)sizes, (uint32_t *)ctx->wg_cnt.p, N, (unsigned long long *)(d_small + 32));
    MP_HIP(hipGetLastError());
    MP_TRY(scan_u32(ctx, (const uint32_t *)ctx->wg_cnt.p, (uint64_t *)ctx->wg_off.p, N));
    hipLaunchKernelGGL(plan_kernel, dim3(1), dim3(64), 0, ctx->stream, (const uint64_t *)ctx->wg_off.p, N, d_small);
    MP_HIP(hipGetLastError());
    uint64_t res[64];
    MP_HIP(hipMemcpyAsync(res, d_small, sizeof(res), hipMemcpyDeviceToHost, ctx->stream));
    if (ctx->batch) {  // plans and statistics per sample
        const int ns = ctx->C;
        uint64_t *d_b = nullptr;
        const size_t words = (size_t)(ns + 1) + (size_t)ns * (MP_NPART + 1) + 2 * (size_t)ns;
        MP_HIP(hipMallocAsync((void **)&d_b, words * sizeof(uint64_t), ctx->stream));
        uint64_t *d_first = d_b, *d_W = d_b + (ns + 1), *d_cs = d_W + (size_t)ns * (MP_NPART + 1);
        MP_HIP(hipMemcpyAsync(d_first, ctx->chunk_ids.data(), sizeof(uint64_t) * (ns + 1), hipMemcpyHostToDevice,
                              ctx->stream));
        MP_TRY(fill_u32(ctx, d_cs, 0, 4 * (uint64_t)ns));
        hipLaunchKernelGGL(sample_plan_kernel, dim3(ns), dim3(64), 0, ctx->stream, (const uint64_t *)ctx->wg_off.p,
                           (const uint64_t *)d_first, d_W);
        hipLaunchKernelGGL(sample_stats_kernel, dim3(nblk(N)), dim3(256), 0, ctx->stream, (const uint32_t *)parent,
                           (const uint32_t *)sizes, (const uint64_t *)d_first, ns, N, (unsigned long long *)d_cs);
        MP_HIP(hipGetLastError());
        ctx->sample_W.resize((size_t)ns * (MP_NPART + 1));
        std::vector<uint64_t> cs(2 * (size_t)ns);
        MP_HIP(hipMemcpyAsync(ctx->sample_W.data(), d_W, sizeof(uint64_t) * ctx->sample_W.size(),
                              hipMemcpyDeviceToHost, ctx->stream));
        MP_HIP(hipMemcpyAsync(cs.data(), d_cs, sizeof(uint64_t) * cs.size(), hipMemcpyDeviceToHost, ctx->stream));
        MP_HIP(hipFreeAsync(d_b, ctx->stream));
        MP_HIP(hipStreamSynchronize(ctx->stream));
        ctx->sample_comp.resize(ns);
        ctx->sample_largest.resize(ns);
        for (int i = 0; i < ns; i++) {
            ctx->sample_comp[i] = cs[2 * i];
            ctx->sample_largest[i] = cs[2 * i + 1];
        }
    }
    MP_HIP(hipFreeAsync(d_small, ctx->stream));
    hipEventRecord(ctx->ev[6], ctx->stream);
    MP_HIP(hipStreamSynchronize(ctx->stream));
    for (int p = 0; p <= MP_NPART; p++) W[p] = res[p];
    if (st) {
        st->pair_unions = unions + res[40];
        st->pair_candidates = n_ordered;
        st->components = res[32];
        st->largest = res[33];
        float ms = 0;
        hipEventElapsedTime(&ms, ctx->ev[5], ctx->ev[6]);
        st->ms_device = ms;
        st->ms_host = (float)host_ms;
    }
    return 0;
}

}  // namespace mp
