// Device-side k-mer generation shared by the enumeration kernels and the
// radix pass that generates its input (canonical k-mer, adj byte, label).
#pragma once
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
    const uint32_t *labels;  // when set: the labels of a scan taken at its start (a two-sweep scan)
    int scan0;
    // when set (a batch, 3-minimizer): each chunk's own upper bucket bound,
    // the one scan of that sample alone would have (chunk of read r from chunk_ids)
    const uint32_t *chunk_hi;
};

enum { kModeHist = 0, kModeCount = 1, kModeEmit = 2, kModeTop = 3 };

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

// Canonical k-mer of the window starting at base b (position p of its read),
// the two streams read through word accessors (memory or a staged copy).
// Returns false if the window holds an 'N'.
template <typename NWord, typename PWord>
__device__ __forceinline__ bool kmer_from(int K, uint64_t b, uint32_t p, NWord nword, PWord pword, uint64_t &canon,
                                          bool &fwd)
{
    {
        const uint64_t w = b >> 6;
        const int o = (int)(b & 63);
        const uint64_t hi = nword(w), lo = nword(w + 1);
        const int s = o + K;
        const uint64_t v = s <= 64 ? (hi >> (64 - s)) : ((hi << (s - 64)) | (lo >> (128 - s)));
        if (v & ((1ull << K) - 1)) return false;
    }
    const uint64_t w = b >> 5;
    const int o = (int)(b & 31);
    const uint64_t hi = pword(w), lo = pword(w + 1);
    const int s = 2 * (o + K);
    uint64_t F = s <= 64 ? (hi >> (64 - s)) : ((hi << (s - 64)) | (lo >> (128 - s)));
    if (K < 32) F &= (1ull << (2 * K)) - 1;
    const uint64_t R = revcomp(F, K);
    if (K == 32 && p > 0) F = 0;
    fwd = F < R;
    canon = fwd ? F : R;
    return true;
}

// A window rolled along a read: forward code F and reverse complement R of
// the last K bases, and the run of bases without 'N' ending at the last one.
// init() builds the window at base b whole; roll() appends base bn (the base
// after the window) and drops the first.  canon() applies the K = 32 quirk of
// kmer_from (forward code 0 past a read's first position).
struct RollKmer {
    uint64_t F, R;
    uint32_t nrun;
    uint64_t pi = ~0ull, ni = ~0ull, pw = 0, nw = 0;  // the 2-bit and 'N' words last read, and their indices
    template <typename NWord, typename PWord>
    __device__ __forceinline__ void init(int K, uint64_t b, NWord nword, PWord pword)
    {
        const int o = (int)(b & 63);
        const uint64_t nh = nword(b >> 6), nl = nword((b >> 6) + 1);
        const int sn = o + K;
        uint64_t v = sn <= 64 ? (nh >> (64 - sn)) : ((nh << (sn - 64)) | (nl >> (128 - sn)));
        v &= K < 64 ? (1ull << K) - 1 : ~0ull;
        nrun = v ? (uint32_t)__builtin_ctzll(v) : (uint32_t)K;  // bases after the window's last 'N'
        const int op = (int)(b & 31);
        const uint64_t ph = pword(b >> 5), pl = pword((b >> 5) + 1);
        const int s = 2 * (op + K);
        F = s <= 64 ? (ph >> (64 - s)) : ((ph << (s - 64)) | (pl >> (128 - s)));
        if (K < 32) F &= (1ull << (2 * K)) - 1;
        R = revcomp(F, K);
    }
    template <typename NWord, typename PWord>
    __device__ __forceinline__ void roll(int K, uint64_t bn, NWord nword, PWord pword)
    {
        if ((bn >> 5) != pi) {  // a word covers 32 (64) bases: most steps reuse it
            pi = bn >> 5;
            pw = pword(pi);
        }
        if ((bn >> 6) != ni) {
            ni = bn >> 6;
            nw = nword(ni);
        }
        const uint64_t c = (pw >> (62 - 2 * (bn & 31))) & 3;
        const uint64_t nb = (nw >> (63 - (bn & 63))) & 1;
        F = (F << 2) | c;
        if (K < 32) F &= (1ull << (2 * K)) - 1;
        R = (R >> 2) | ((3 - c) << (2 * K - 2));
        nrun = nb ? 0u : nrun + 1;
    }
    __device__ __forceinline__ bool valid(int K) const { return nrun >= (uint32_t)K; }
    __device__ __forceinline__ uint64_t canon(int K, uint32_t p) const
    {
        const uint64_t f = (K == 32 && p > 0) ? 0 : F;
        return f < R ? f : R;
    }
};

// Calls fn(canonical k-mer, q) for every window q without 'N' of a group of nr
// reads (window prefix pref[0..nr], first bases base[0..nr)): each of the
// nlanes lanes takes a contiguous run of the group's windows and rolls the
// k-mer from one window to the next, building a read's first window whole.
template <typename NWord, typename PWord, typename Fn>
__device__ __forceinline__ void for_group_windows(int K, const uint32_t *pref, const uint64_t *base, int nr,
                                                  uint32_t lane_id, uint32_t nlanes, NWord nword, PWord pword, Fn fn)
{
    const uint32_t total = pref[nr];
    const uint32_t cs = (total + nlanes - 1) / nlanes;
    uint32_t q = lane_id * cs;
    const uint32_t qe = min(q + cs, total);
    if (q >= qe) return;
    int lo = 0, h = nr - 1;  // the read of q: last lo with pref[lo] <= q
    while (lo < h) {
        const int mid = (lo + h + 1) >> 1;
        if (pref[mid] <= q) lo = mid; else h = mid - 1;
    }
    uint32_t p = q - pref[lo], pend = pref[lo + 1] - pref[lo];
    RollKmer k;
    k.init(K, base[lo] + p, nword, pword);
    for (;;) {
        if (k.valid(K)) fn(k.canon(K, p), q);
        if (++q >= qe) break;
        if (++p < pend) {
            k.roll(K, base[lo] + p + K - 1, nword, pword);
        } else {  // the next read with a window (q < total: there is one)
            do lo++; while (pref[lo + 1] == pref[lo]);
            p = 0;
            pend = pref[lo + 1] - pref[lo];
            k.init(K, base[lo], nword, pword);
        }
    }
}

// adj byte from the bytes before and after the window (enumerateKmers.c:215, 239)
__device__ __forceinline__ uint8_t adj_from(uint8_t prev, uint8_t next, bool fwd)
{
    return fwd ? adj_of(prev, next) : adj_of(rc_char(next), rc_char(prev));
}

// label of read r: the read itself in the first scan; later scans label a
// read with its component root, or with itself once the filter marked it
__device__ __forceinline__ uint32_t label_of(const KParams &P, uint64_t r)
{
    if (P.labels) return P.labels[r];
    return P.scan0 ? (uint32_t)r : (P.bad[r] ? (uint32_t)r : P.parent[r]);
}

struct GenSource {
    KParams P;                                // reads, K, bucket range, labels
    uint64_t h[kSortPasses * kSortRadix];     // digit counts of the tuples
    uint64_t npos;                            // k-mer windows of all reads
    // when set: digit counts of radix pass gcnt_pass per generating-pass tile
    // ([tile][digit], from the m-mer histogram), so that pass needs no look-back
    const uint32_t *gcnt = nullptr;
    uint64_t gcnt_tiles = 0;
    int gcnt_pass = -1;
    bool gcnt_seg = false;  // tiles of every chunk's windows on their own (a batch's samples)
    std::vector<uint64_t> seg_tuples;  // with gcnt_seg: the k-mers of every chunk
};

}  // namespace mp
