// Internal declarations of the metapart_amd HIP library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "../../include/metapart_amd.h"

namespace mp {

constexpr int kWave = 64;
constexpr uint32_t kBad = 1;  // bit of bad[]: label marked by the frequency filter
constexpr int kSortRadix = 256;
constexpr uint64_t kStageBytes = 256ull << 20;  // each of a ctx's two pinned host staging buffers
constexpr int kSortPasses = 10;  // radix passes: adj, k-mer bytes 0..7, sample of the label

// ---------------------------------------------------------------- device helpers
__device__ __forceinline__ uint64_t ld_agent(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_agent(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t lanemask_lt()
{
    const unsigned lane = __lane_id();
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// Adds v summed over the workgroup to *ctr with one atomic (blockDim.x a
// multiple of 64, at most 1024; every lane of the workgroup must call it; it
// synchronises the workgroup, so calls may follow each other).
__device__ __forceinline__ void block_add(uint32_t v, unsigned long long *ctr)
{
    __shared__ uint32_t s_part[16];
    __syncthreads();  // a previous call may still be reading s_part
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_down(v, d, kWave);
    if (__lane_id() == 0) s_part[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (unsigned w = 0; w < blockDim.x / kWave; w++) t += s_part[w];
        if (t) atomicAdd(ctr, (unsigned long long)t);
    }
}
// grid of at most 2048 workgroups of 256 lanes for a grid-stride loop over n items
inline unsigned stride_grid(uint64_t n)
{
    const uint64_t b = (n + 255) / 256;
    return (unsigned)(b < 2048 ? (b ? b : 1) : 2048);
}
// workgroups of 256 lanes for a wave-per-item grid-stride loop over n items:
// one wave per item up to 2^20 workgroups (a dispatch holds < 2^32 lanes)
inline unsigned wave_grid(uint64_t n)
{
    const uint64_t b = (n + 3) / 4;
    return (unsigned)(b < (1u << 20) ? (b ? b : 1) : (1u << 20));
}
// a dispatch holds fewer than 2^32 work-items (the packet's grid size is 32-bit)
inline bool grid_fits(uint64_t blocks, uint64_t lanes) { return blocks * lanes < (1ull << 32); }

// Reverse-complement of a 2-bit packed k-mer (first base most significant,
// A0 C1 G2 T3, complement = 3 - code): reverse the 2-bit groups of ~x.
__device__ __forceinline__ uint64_t revcomp(uint64_t x, int K)
{
    x = ~x;
    x = __builtin_bitreverse64(x);
    x = ((x >> 1) & 0x5555555555555555ull) | ((x & 0x5555555555555555ull) << 1);
    return x >> (64 - 2 * K);
}

// ---------------------------------------------------------------- context
struct DevArray {
    void *p = nullptr;
    uint64_t bytes = 0;
};

}  // namespace mp

struct mp_ctx {
    mp_cfg cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8]{};
    hipEvent_t pev[24]{};  // per radix pass (timing of the dominant kernel)
    hipEvent_t cev[mp::kSortPasses + 2]{};  // per counted radix pass: before its digit count (last: generating pass)
    float sort_pass_avg = 0, sort_hist_ms = 0;
    float sort_count_ms = 0;   // digit counts and offsets of the counted passes of the last sort
    float sort_gen_ms = 0;     // the generating radix pass of the last sort (0: none)
    float local_ms = 0;        // local rank pass of the last sort
    uint64_t local_long = 0;   // groups it handed to the long-group kernel
    uint64_t local_moved = 0;  // tuples it moved (the rest stayed in place)
    std::string err;
    uint64_t budget = 0, held = 0;
    // arena mode (the allocation-free entry point): every device buffer is
    // carved from [arena, arena + arena_bytes), nothing is allocated or freed
    uint8_t *arena = nullptr;
    uint64_t arena_bytes = 0, arena_used = 0;
    uint16_t epoch = 0;  // look-back status epoch of the radix sort
    int status_fmt = 0;  // 1: 32-bit status words, 2: 64-bit (the array is cleared on a change)

    // reads
    const uint8_t *d_fq = nullptr;  // FASTQ bytes (owned if fq_owned)
    bool fq_owned = false;
    uint64_t fq_bytes = 0;
    uint64_t N = 0;  // reads
    int C = 0;       // chunks
    uint64_t total_bases = 0;

    // per-read arrays
    mp::DevArray fq, rec_off, seq_off, seq_len, base_off, pos_off, packed, nmask;
    // graph state
    mp::DevArray parent, bad, filtered, sizes;
    mp::DevArray gparent, gsize;  // pair step: components of the candidate-pair graph
    mp::DevArray pairc;           // pair step: compact ids of the ordered pass
    // scan buffers (ping-pong)
    mp::DevArray keys[2], lab[2], adj[2], zbits;
    mp::DevArray runbits;  // filter -> union: starts of runs to join, starts of all runs
    mp::DevArray dig, rts; // counted radix passes: next pass's digit per tuple; per-tile counts and offsets
    mp::DevArray gcnt;     // m-mer histogram: digit counts per generating-pass tile (a scan over every k-mer)
    mp::DevArray segtab;   // a batch's per-sample generating-pass tiles: first tile, first window
    mp::DevArray sortaux;  // radix sort: digit bases, segment tables, generating-pass tile table
    // captured radix pass chains (mp_cfg.sort_graph), the latest last, with the
    // shape and buffers each was captured for (a batch sorts its tuples, the
    // pair step's pairs and the writer's items: each chain keeps its graph)
    std::vector<std::pair<std::vector<uint64_t>, hipGraphExec_t>> sort_graphs;
    uint64_t sort_graph_captures = 0, sort_graph_replays = 0;
    int sort_graph_last = 0;  // the last sort's chain: 0 launched pass by pass, 1 captured, 2 replayed
    std::vector<uint64_t> seg_wpos;  // first window of every sample (C + 1)
    int khist_copies = 32; // copies of the byte counts in khist (one per sample for a batch)
    bool gcnt_seg = false; // gcnt over every sample's own tiles
    bool gcnt_valid = false;
    uint64_t gcnt_tile = 0, gcnt_tiles = 0;
    int gcnt_byte = 0;
    mp::DevArray wg_cnt, wg_off, status, scratch, hist;
    mp::DevArray scanst;  // look-back status of the exclusive scans
    mp::DevArray hpart;   // m-mer histogram: tile table, partition bounds, work list
    mp::DevArray labsnap;  // a two-sweep scan's labels, taken at its start
    bool use_labsnap = false;
    mp::DevArray khist;   // m-mer histogram: counts of the k-mer bytes the scans' radix passes read
    bool khist_valid = false;
    uint32_t khist_mask = 0;  // passes (bit 1 + byte) khist counts
    uint64_t kmer_total = 0;  // k-mers of the loaded reads (last m-mer histogram)
    // writer
    mp::DevArray out, items[2], ilens[2], iaux[2];
    // pinned host staging of large uploads (two buffers, one copy in flight)
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    bool stage_used[2] = {false, false};
    // the ordered pair step's host side: its pair list, sizes, component ranges
    // and final roots in pinned memory (copied at the link's rate, no page
    // faults after the first call), and its union-find nodes
    void *pair_host = nullptr;
    uint64_t pair_host_bytes = 0;
    std::vector<uint64_t> pair_nodes;
    std::vector<uint64_t> chunk_ofs;  // byte offset of each chunk (C+1)
    int cur = 0;               // buffer holding the last scan's sorted tuples
    uint64_t last_n = 0;       // tuples of the last scan
    bool have_scan = false;
    bool finished = false;
    int scans_done = 0;
    std::vector<uint64_t> chunk_ids;  // first read id of each chunk (C+1)
    // batch of independent samples: one chunk per sample
    bool batch = false;
    std::vector<uint64_t> sample_W;        // 17 per sample, local read ids
    std::vector<uint64_t> sample_comp;     // components per sample
    std::vector<uint64_t> sample_largest;  // largest component per sample
    std::vector<uint64_t> sample_tuples;   // tuple offsets of the samples (C+1)

    // every device buffer the ctx holds
    template <typename F> void for_each_array(F f)
    {
        mp::DevArray *arrs[] = {&fq, &rec_off, &seq_off, &seq_len, &base_off, &pos_off, &packed, &nmask, &parent, &bad,
                                &filtered, &sizes, &gparent, &gsize, &pairc, &scanst, &hpart, &khist, &labsnap,
                                &runbits, &dig, &rts, &gcnt, &segtab, &sortaux, &keys[0], &keys[1], &lab[0], &lab[1], &adj[0], &adj[1], &zbits, &wg_cnt,
                                &wg_off, &status, &scratch, &hist, &out, &items[0], &items[1], &ilens[0], &ilens[1],
                                &iaux[0], &iaux[1]};
        for (mp::DevArray *a : arrs) f(*a);
    }

    int fail(int code, const char *fmt, ...)
    {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
};

namespace mp {

#define MP_HIP(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return ctx->fail(MP_EDEVICE, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)
#define MP_TRY(x)              \
    do {                       \
        int r_ = (x);          \
        if (r_) return r_;     \
    } while (0)

// Grow-only device allocation charged against the ctx budget.
int ensure(mp_ctx *ctx, DevArray &a, uint64_t bytes);
void release(mp_ctx *ctx, DevArray &a);
// Stream-ordered temporary device memory (hipMallocAsync / hipFreeAsync, or
// the ctx's arena in arena mode).
int tmp_alloc(mp_ctx *ctx, void **p, uint64_t bytes);
int tmp_free(mp_ctx *ctx, void *p);

// Exclusive scan of n u32 or u64 values into out (n+1 entries: out[n] = total).
int scan_u32(mp_ctx *ctx, const uint32_t *in, uint64_t *out, uint64_t n);
int scan_u64(mp_ctx *ctx, const uint64_t *in, uint64_t *out, uint64_t n);
// Fill with a 32-bit pattern via a kernel (safe under graph capture).
int fill_u32(mp_ctx *ctx, void *p, uint32_t v, uint64_t n_words);

// stages (each .hip file)
int parse_reads(mp_ctx *ctx, const uint64_t *chunk_ofs, const uint64_t *chunk_first_id, int C);
int count_chunks_device(mp_ctx *ctx, const uint64_t *chunk_first_id, int C, uint32_t *d_counts);
// A batch (chunk = sample): the byte counts of every sample and the generating
// pass's digit counts over tiles of every sample's windows (gen_histograms
// then gives per-sample tuple counts for a segmented sort).
int batch_counts(mp_ctx *ctx);
// per-bucket sums over chunks (u64 per bucket) of a C x (2^20+2) count table
int bucket_totals(mp_ctx *ctx, const uint32_t *d_counts, int C, uint64_t *d_tot);
// 3-minimizer k-mers of a scan: ordered enumeration into the tuple buffer 0
int enumerate_scan(mp_ctx *ctx, int s, uint64_t lo, uint64_t hi, uint64_t *n_out, bool per_chunk_hi = false);
// The scan's tuples generated by the first radix pass instead of an
// enumeration: the reads' k-mers in [lo, hi) and their digit histograms.
struct GenSource;
// passes whose digits the sort of a scan's tuples reads (bit p = radix pass p)
uint32_t scan_pass_mask(int K);
// The k-mer digits of the radix passes: pass 1 + b takes bits [shift[b],
// shift[b] + bits[b]) of the key.  Bytes by default.  A scan's layout may
// split the k-mer's bits above bit 48 evenly over the last two passes (k = 31:
// 7 + 7 bits instead of 8 + 6): fewer digits per pass keep a tile's digit runs
// longer, and its partial cache lines fewer.
struct KeyLayout {
    uint8_t shift[8], bits[8];
    __host__ __device__ __forceinline__ uint32_t digit(uint64_t k, int b) const
    {
        return (uint32_t)(k >> shift[b]) & ((1u << bits[b]) - 1u);
    }
};
KeyLayout byte_layout();
// the layout of the ctx's scans (cfg.sort_digit_bits: 8 bytes; 0 or 7 the split top)
KeyLayout scan_layout(const mp_ctx *ctx);
// kSortPasses x kSortRadix digit counts of the scan's tuples (host, passes of
// pass_mask) and their number; fills *g for radix_sort_tuples.  all: the scan
// holds every k-mer, so the counts of the last m-mer histogram serve.
int gen_histograms(mp_ctx *ctx, int s, uint64_t lo, uint64_t hi, uint32_t pass_mask, bool all, GenSource *g,
                   uint64_t *n_out);
// Samples of a batch sorted each on its own (host arrays of ns + 1 entries):
// sample s's tuples end up in [tpos[s], tpos[s+1]); the generating pass takes
// its windows [wpos[s], wpos[s+1]) (and the m-mer histogram's per-tile digit
// counts over tiles of each sample's windows).
struct SortSegments {
    int ns = 0;
    const uint64_t *tpos = nullptr;
    const uint64_t *wpos = nullptr;
};
int radix_sort_tuples(mp_ctx *ctx, uint64_t n, int *passes_out, const GenSource *gen = nullptr,
                      const SortSegments *segs = nullptr);
// tuples (k-mer windows for the generating pass) per radix tile of the ctx's tile shape
uint64_t sort_tile_items(const mp_ctx *ctx);
// Stable LSD passes over (keys, lab, adj) ping-pong buffers; pass p of pass_mask:
// 0 = adj byte, 1..8 = key byte p-1, 9 = sample of the label (d_grp_first:
// n_grp + 1 first read ids on the device; grp_base, if given, receives the
// n_grp + 1 tuple offsets of the samples).  *cur selects the buffer holding the data.
// With gen, the first pass generates the tuples (gen's histograms replace the
// histogram read) into the buffer *cur does not select.
// Key passes take their digits from *kl (nullptr: bytes).
int radix_sort_generic(mp_ctx *ctx, DevArray *keys, DevArray *lab, DevArray *adj, int *cur, uint64_t n,
                       uint32_t pass_mask, int *passes_out, const uint64_t *d_grp_first, int n_grp,
                       uint64_t *grp_base, const GenSource *gen = nullptr, const SortSegments *segs = nullptr,
                       const KeyLayout *kl = nullptr);
int iota_u32(mp_ctx *ctx, uint32_t *p, uint64_t n);
int filter_union_scan(mp_ctx *ctx, uint64_t n, const uint64_t *seg_bounds, int T, mp_scan_stats *st);
// Same over segments at the given tuple positions (nseg + 1, host) with
// per-segment quirk groups (2 per segment: first and end tuple, host).
int filter_union_groups(mp_ctx *ctx, uint64_t n, const uint64_t *seg_pos, int nseg, const uint64_t *grp,
                        mp_scan_stats *st);
// A scan over bucket parts (cuts[0..P], bucket boundaries) that together do
// not fit the budget: filter sweep, then union sweep (see mp_graph.hip).
// make_part(lo, hi, &n) leaves the part's sorted tuples in the current buffers.
int two_sweep_scan(mp_ctx *ctx, const std::vector<uint64_t> &cuts, const uint64_t *seg_bounds, int T,
                   const std::function<int(uint64_t, uint64_t, uint64_t *)> &make_part, mp_scan_stats *st);
// Per-bucket k-mer counts of [lo, hi) (host prefix, hi - lo + 1 entries), by
// device atomics: no staging memory beyond the counts.
int bucket_prefix_range(mp_ctx *ctx, uint64_t lo, uint64_t hi, std::vector<uint64_t> &pre);
// Labels of scan s for every read as they stand now; the scan's k-mer
// kernels use them until use_labsnap is cleared.
int snapshot_labels(mp_ctx *ctx, int s);
int finish_components(mp_ctx *ctx, uint64_t *W, mp_finish_stats *st);
int write_partitions(mp_ctx *ctx, const char *out_prefix, int T, const uint64_t *W, const int64_t *il_start,
                     uint64_t *bytes_out);

}  // namespace mp
