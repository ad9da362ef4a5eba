// C-ABI entry points of libmetapart_amd (declared in include/metapart_amd.h).
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "mp_internal.h"
#include "mp_kmer.h"

using namespace mp;

extern "C" {

int mp_ctx_create(const mp_cfg *cfg, mp_ctx **out)
{
    if (!cfg || !out) return MP_EINVAL;
    *out = nullptr;
    if (cfg->k <= 10 || cfg->k > 32) return MP_EINVAL;
    // 8: byte digits; 0 or 7: the scans split the k-mer's top bits over their last two passes
    if (cfg->sort_digit_bits != 0 && cfg->sort_digit_bits != 7 && cfg->sort_digit_bits != 8) return MP_EINVAL;
    if (cfg->sort_tile != 0 && cfg->sort_tile != 3072 && cfg->sort_tile != 4096 && cfg->sort_tile != 6144 &&
        cfg->sort_tile != 8192)
        return MP_EINVAL;
    if (cfg->sort_graph != 0 && cfg->sort_graph != 1) return MP_EINVAL;
    mp_ctx *ctx = new (std::nothrow) mp_ctx();
    if (!ctx) return MP_ENOMEM;
    ctx->cfg = *cfg;
    ctx->device = cfg->device;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0) {
        (void)hipGetLastError();
        delete ctx;
        return MP_EDEVICE;
    }
    if (hipSetDevice(cfg->device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return MP_EDEVICE;
    }
    for (auto &e : ctx->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete ctx;
            return MP_EDEVICE;
        }
    for (auto &e : ctx->pev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete ctx;
            return MP_EDEVICE;
        }
    for (auto &e : ctx->cev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete ctx;
            return MP_EDEVICE;
        }
    size_t fre = 0, tot = 0;
    (void)hipMemGetInfo(&fre, &tot);
    ctx->budget = cfg->hbm_budget ? cfg->hbm_budget : (uint64_t)(fre * 0.9);
    *out = ctx;
    return 0;
}

int mp_ctx_destroy(mp_ctx *ctx)
{
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->for_each_array([&](DevArray &a) { release(ctx, a); });
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : ctx->pev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : ctx->cev)
        if (e) (void)hipEventDestroy(e);
    for (auto &g : ctx->sort_graphs) (void)hipGraphExecDestroy(g.second);
    for (int b = 0; b < 2; b++) {
        if (ctx->stage_ev[b]) (void)hipEventDestroy(ctx->stage_ev[b]);
        if (ctx->stage[b]) (void)hipHostFree(ctx->stage[b]);
    }
    if (ctx->pair_host) (void)hipHostFree(ctx->pair_host);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

const char *mp_last_error(const mp_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }
uint64_t mp_ctx_device_bytes(const mp_ctx *ctx) { return ctx ? ctx->held : 0; }

static int load_common(mp_ctx *ctx, const uint64_t *chunk_ofs, const uint64_t *chunk_first_id, int C)
{
    if (C <= 0 || !chunk_ofs) return ctx->fail(MP_EINVAL, "need at least one chunk");
    if (chunk_ofs[0] != 0 || chunk_ofs[C] != ctx->fq_bytes)
        return ctx->fail(MP_EINVAL, "chunks must cover the FASTQ bytes exactly");
    for (int c = 0; c < C; c++)
        if (chunk_ofs[c + 1] < chunk_ofs[c]) return ctx->fail(MP_EINVAL, "chunk offsets must not decrease");
    ctx->have_scan = false;
    ctx->finished = false;
    ctx->scans_done = 0;
    ctx->last_n = 0;
    ctx->batch = false;
    ctx->sample_W.clear();
    ctx->khist_valid = false;
    MP_TRY(parse_reads(ctx, chunk_ofs, chunk_first_id, C));
    ctx->chunk_ofs.assign(chunk_ofs, chunk_ofs + C + 1);
    const uint64_t N = ctx->N;
    MP_TRY(ensure(ctx, ctx->parent, N * sizeof(uint32_t)));
    MP_TRY(ensure(ctx, ctx->bad, N + 16));
    MP_TRY(ensure(ctx, ctx->filtered, N + 16));
    MP_TRY(ensure(ctx, ctx->sizes, N * sizeof(uint32_t)));
    return 0;
}

// Host to device copy of FASTQ bytes.  Large inputs go through two pinned
// staging buffers: host threads fill one while the DMA engine drains the
// other, so pageable memory is read once at memory speed and PCIe runs at the
// pinned rate.
constexpr int kStageThreads = 8;

static void parallel_copy(char *dst, const char *src, uint64_t n)
{
    std::vector<std::thread> th;
    const uint64_t part = (n + kStageThreads - 1) / kStageThreads;
    for (int t = 0; t < kStageThreads; t++) {
        const uint64_t a = t * part, b = std::min(n, a + part);
        if (a >= b) break;
        th.emplace_back([=] { memcpy(dst + a, src + a, b - a); });
    }
    for (auto &x : th) x.join();
}

static int copy_to_device(mp_ctx *ctx, uint8_t *dst, const char *src, uint64_t nbytes)
{
    if (nbytes < 2 * kStageBytes) {
        MP_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyHostToDevice, ctx->stream));
        return 0;
    }
    for (int b = 0; b < 2; b++) {
        if (!ctx->stage[b]) MP_HIP(hipHostMalloc(&ctx->stage[b], kStageBytes, hipHostMallocDefault));
        if (!ctx->stage_ev[b]) MP_HIP(hipEventCreateWithFlags(&ctx->stage_ev[b], hipEventDisableTiming));
    }
    for (uint64_t off = 0, c = 0; off < nbytes; off += kStageBytes, c++) {
        const int b = (int)(c & 1);
        const uint64_t len = std::min(kStageBytes, nbytes - off);
        if (ctx->stage_used[b]) MP_HIP(hipEventSynchronize(ctx->stage_ev[b]));  // its previous copy has left
        parallel_copy((char *)ctx->stage[b], src + off, len);
        MP_HIP(hipMemcpyAsync(dst + off, ctx->stage[b], len, hipMemcpyHostToDevice, ctx->stream));
        MP_HIP(hipEventRecord(ctx->stage_ev[b], ctx->stream));
        ctx->stage_used[b] = true;
    }
    return 0;
}

static int upload_fq(mp_ctx *ctx, const char *fq, uint64_t nbytes)
{
    MP_TRY(ensure(ctx, ctx->fq, nbytes + 64));
    MP_TRY(copy_to_device(ctx, (uint8_t *)ctx->fq.p, fq, nbytes));
    MP_TRY(fill_u32(ctx, (uint8_t *)ctx->fq.p + ((nbytes + 3) & ~3ull), 0, 8));
    ctx->d_fq = (const uint8_t *)ctx->fq.p;
    ctx->fq_owned = true;
    ctx->fq_bytes = nbytes;
    return 0;
}

int mp_count_chunks(mp_ctx *ctx, const char *fq, uint64_t nbytes, const uint64_t *chunk_ofs, int C, uint32_t *counts,
                    uint64_t *reads_per_chunk, uint64_t *total_kmers)
{
    if (!ctx || !fq || !counts) return MP_EINVAL;
    MP_HIP(hipSetDevice(ctx->device));
    MP_TRY(upload_fq(ctx, fq, nbytes));
    MP_TRY(load_common(ctx, chunk_ofs, nullptr, C));
    const uint64_t words = (uint64_t)C * (MP_NBUCKETS + 2);
    MP_TRY(ensure(ctx, ctx->scratch, words * sizeof(uint32_t) + 4096));
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

int mp_count_loaded(mp_ctx *ctx, uint32_t *counts, uint64_t *hist, uint64_t *reads_per_chunk, uint64_t *total_kmers)
{
    if (!ctx) return MP_EINVAL;
    if (ctx->N == 0) return ctx->fail(MP_ESTATE, "mp_count_loaded before mp_load_reads");
    MP_HIP(hipSetDevice(ctx->device));
    const int C = ctx->C;
    const uint64_t words = (uint64_t)C * (MP_NBUCKETS + 2);
    MP_TRY(ensure(ctx, ctx->scratch, words * sizeof(uint32_t) + 4096));
    MP_TRY(ensure(ctx, ctx->hist, (2 * (uint64_t)MP_NBUCKETS + 2) * sizeof(uint64_t)));
    uint32_t *d_counts = (uint32_t *)ctx->scratch.p;
    MP_TRY(count_chunks_device(ctx, ctx->chunk_ids.data(), C, d_counts));
    uint64_t *d_tot = (uint64_t *)ctx->hist.p, *d_pre = d_tot + MP_NBUCKETS;
    MP_TRY(bucket_totals(ctx, d_counts, C, d_tot));
    if (counts) MP_HIP(hipMemcpyAsync(counts, d_counts, words * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    MP_TRY(scan_u64(ctx, d_tot, d_pre, MP_NBUCKETS));
    // straight into the caller's buffers
    uint64_t total = 0;
    if (hist)
        MP_HIP(hipMemcpyAsync(hist, d_pre, sizeof(uint64_t) * (MP_NBUCKETS + 1), hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipMemcpyAsync(&total, d_pre + MP_NBUCKETS, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    if (total_kmers) *total_kmers = total;
    if (reads_per_chunk)
        for (int c = 0; c < C; c++) reads_per_chunk[c] = ctx->chunk_ids[c + 1] - ctx->chunk_ids[c];
    return 0;
}

int mp_chunk_first_ids(mp_ctx *ctx, uint64_t *ids)
{
    if (!ctx || !ids) return MP_EINVAL;
    if (ctx->N == 0) return ctx->fail(MP_ESTATE, "no reads loaded");
    for (int c = 0; c <= ctx->C; c++) ids[c] = ctx->chunk_ids[c];
    return 0;
}

int mp_load_reads(mp_ctx *ctx, const char *fq, uint64_t nbytes, const uint64_t *chunk_ofs,
                  const uint64_t *chunk_first_id, int C)
{
    if (!ctx || !fq) return MP_EINVAL;
    MP_HIP(hipSetDevice(ctx->device));
    MP_TRY(upload_fq(ctx, fq, nbytes));
    return load_common(ctx, chunk_ofs, chunk_first_id, C);
}

int mp_load_reads_device(mp_ctx *ctx, const void *d_fq, uint64_t nbytes, const uint64_t *chunk_ofs,
                         const uint64_t *chunk_first_id, int C)
{
    if (!ctx || !d_fq) return MP_EINVAL;
    MP_HIP(hipSetDevice(ctx->device));
    // a FASTQ buffer an earlier host load uploaded is no longer needed: it
    // would count against the budget of every later scan
    if (ctx->fq_owned) {
        MP_HIP(hipStreamSynchronize(ctx->stream));
        release(ctx, ctx->fq);
    }
    ctx->d_fq = (const uint8_t *)d_fq;
    ctx->fq_owned = false;
    ctx->fq_bytes = nbytes;
    return load_common(ctx, chunk_ofs, chunk_first_id, C);
}

// Device bytes a scan holds per tuple: two tuple buffers (13 B each), the
// zeroed-run and run bitmaps, the look-back status, the counted passes'
// digit bytes and count tables, and the growth slack of the buffers.
constexpr uint64_t kScanBytesPerTuple = 31;

// Tuples a scan can hold within the budget: the bytes not held by anything
// else, less a label snapshot and the room the pair step and the writer will
// need, over kScanBytesPerTuple (MP_SCAN_PART_TUPLES lowers it, for tests).
static uint64_t scan_tuple_room(mp_ctx *ctx)
{
    const uint64_t N = ctx->N;
    const uint64_t tuple_held = ctx->keys[0].bytes + ctx->keys[1].bytes + ctx->lab[0].bytes + ctx->lab[1].bytes +
                                ctx->adj[0].bytes + ctx->adj[1].bytes + ctx->zbits.bytes + ctx->runbits.bytes +
                                ctx->dig.bytes + ctx->rts.bytes +
                                ctx->status.bytes;
    const uint64_t other = ctx->held - tuple_held;
    const uint64_t label_bytes = ctx->labsnap.bytes ? 0 : N * sizeof(uint32_t) + 64;
    const uint64_t later_held = ctx->gparent.bytes + ctx->gsize.bytes + ctx->pairc.bytes + ctx->items[0].bytes +
                                ctx->items[1].bytes + ctx->ilens[0].bytes + ctx->ilens[1].bytes + ctx->iaux[0].bytes +
                                ctx->iaux[1].bytes + ctx->out.bytes;
    // (measured: 48 B per read for the pair step, 66 B per read of items plus
    // the output for the writer; with the buffers' growth slack of 1/16)
    const uint64_t later_need = (128 * N + ctx->fq_bytes) / 16 * 17 + (1u << 20);
    const uint64_t reserve = later_need > later_held ? later_need - later_held : 0;
    // the generating pass's look-back status covers every k-mer window, not only the scan's
    const uint64_t npos_status = ctx->total_bases / 16;
    const uint64_t fixed = other + label_bytes + reserve + npos_status;
    const uint64_t avail = ctx->budget > fixed ? ctx->budget - fixed : 0;
    uint64_t cap = avail / kScanBytesPerTuple;
    if (const char *e = getenv("MP_SCAN_PART_TUPLES")) cap = std::min<uint64_t>(cap, strtoull(e, nullptr, 10));
    return cap;
}

// The two-sweep scan: bucket parts of at most cap tuples, cut where the
// per-bucket counts allow; the labels are fixed at the scan's start.
static int scan_two_sweeps(mp_ctx *ctx, int s, uint64_t lo, uint64_t hi, const uint64_t *seg_bounds, int T,
                           uint64_t expected, uint64_t cap, mp_scan_stats *st)
{
    if (cap == 0) return ctx->fail(MP_ENOMEM, "the device budget leaves no room for a scan's tuples");
    if (ctx->cfg.min3)  // the window state machine emits a different set over a bucket sub-range (C.9)
        return ctx->fail(MP_ENOMEM, "scan %d needs %llu tuples, more than the budget's %llu, and 3-minimizer scans "
                                    "cannot be split into bucket parts", s, (unsigned long long)expected,
                         (unsigned long long)cap);
    std::vector<uint64_t> pre;
    MP_TRY(bucket_prefix_range(ctx, lo, hi, pre));
    if (pre.back() != expected)
        return ctx->fail(MP_ECOUNT, "scan %d holds %llu tuples, the histogram sized %llu", s,
                         (unsigned long long)pre.back(), (unsigned long long)expected);
    struct Snap {  // the snapshot serves this scan only, whatever way it ends
        mp_ctx *c;
        ~Snap()
        {
            c->use_labsnap = false;
            release(c, c->labsnap);
        }
    } snap{ctx};
    // tuple buffers from earlier steps may be larger than a part needs (the
    // histogram's staging): let each part size them
    MP_HIP(hipStreamSynchronize(ctx->stream));
    for (DevArray *a : {&ctx->keys[0], &ctx->keys[1], &ctx->lab[0], &ctx->lab[1], &ctx->adj[0], &ctx->adj[1],
                        &ctx->zbits, &ctx->runbits, &ctx->dig, &ctx->rts})
        release(ctx, *a);
    MP_TRY(snapshot_labels(ctx, s));
    cap = std::min(cap, scan_tuple_room(ctx));  // the counts and the snapshot took their share
    if (cap == 0) return ctx->fail(MP_ENOMEM, "the device budget leaves no room for a scan's tuples");
    std::vector<uint64_t> cuts{lo};
    while (cuts.back() < hi) {
        const uint64_t a = cuts.back() - lo;
        // last bucket end e with pre[e] - pre[a] <= cap
        const uint64_t e = (uint64_t)(std::upper_bound(pre.begin() + a + 1, pre.end(), pre[a] + cap) - pre.begin()) - 1;
        if (e == a)
            return ctx->fail(MP_ENOMEM, "bucket %llu alone holds %llu tuples, more than the budget's %llu",
                             (unsigned long long)(lo + a), (unsigned long long)(pre[a + 1] - pre[a]),
                             (unsigned long long)cap);
        cuts.push_back(lo + e);
    }
    const uint32_t mask = scan_pass_mask(ctx->cfg.k);
    int passes = 0;
    auto make_part = [&](uint64_t plo, uint64_t phi, uint64_t *n) -> int {
        std::unique_ptr<GenSource> gen;
        if (ctx->cfg.min3) {
            MP_TRY(enumerate_scan(ctx, s, plo, phi, n));
        } else {
            gen.reset(new GenSource());
            MP_TRY(gen_histograms(ctx, s, plo, phi, mask, false, gen.get(), n));
        }
        const uint64_t want = pre[phi - lo] - pre[plo - lo];
        if (*n != want)
            return ctx->fail(MP_ECOUNT, "scan %d part [%llu, %llu) made %llu tuples, counted %llu", s,
                             (unsigned long long)plo, (unsigned long long)phi, (unsigned long long)*n,
                             (unsigned long long)want);
        return radix_sort_tuples(ctx, *n, &passes, gen.get());
    };
    const auto t0 = std::chrono::steady_clock::now();
    MP_TRY(two_sweep_scan(ctx, cuts, seg_bounds, T, make_part, st));
    st->sort_passes = passes;
    st->ms_sort = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() -
                  st->ms_filter - st->ms_union;
    return 0;
}

int mp_scan(mp_ctx *ctx, int s, uint64_t lo, uint64_t hi, const uint64_t *seg_bounds, int T, uint64_t expected_tuples,
            mp_scan_stats *stats)
{
    if (!ctx || !seg_bounds || T <= 0) return MP_EINVAL;
    if (ctx->N == 0) return ctx->fail(MP_ESTATE, "mp_scan before mp_load_reads");
    // the reference's genHistograms counts an odd number of reads, but its
    // pair step reads id N of an odd count and its metapart crashes there
    if (ctx->N & 1)
        return ctx->fail(MP_EINVAL, "odd read count %llu (the pair step reads id N)", (unsigned long long)ctx->N);
    if (s != ctx->scans_done) return ctx->fail(MP_ESTATE, "scan %d out of order (next is %d)", s, ctx->scans_done);
    if (seg_bounds[0] != lo || seg_bounds[T] != hi) return ctx->fail(MP_EINVAL, "segments must span [lo, hi)");
    for (int t = 0; t < T; t++)
        if (!(seg_bounds[t] < seg_bounds[t + 1])) return ctx->fail(MP_EPLAN, "empty or unordered thread range %d", t);
    MP_HIP(hipSetDevice(ctx->device));
    mp_scan_stats st;
    memset(&st, 0, sizeof(st));
    const uint64_t N = ctx->N;
    if (s == 0) {  // metapart.c:3933-3954
        MP_TRY(iota_u32(ctx, (uint32_t *)ctx->parent.p, N));
        MP_TRY(fill_u32(ctx, ctx->bad.p, 0, (N + 3) / 4));
        MP_TRY(fill_u32(ctx, ctx->filtered.p, 0, (N + 3) / 4));
    }
    // a scan whose tuples do not fit the budget goes twice over bucket parts that do
    {
        const uint64_t cap = scan_tuple_room(ctx);
        if (expected_tuples && expected_tuples > cap) {
            MP_TRY(scan_two_sweeps(ctx, s, lo, hi, seg_bounds, T, expected_tuples, cap, &st));
            if (stats) *stats = st;
            ctx->last_n = 0;
            ctx->have_scan = false;  // only the last part's tuples are resident: nothing to dump
            ctx->scans_done++;
            return 0;
        }
    }
    hipEventRecord(ctx->ev[0], ctx->stream);
    uint64_t n = 0;
    // every k-mer: the first radix pass generates the tuples from the reads;
    // 3-minimizer k-mers (a per-read window state machine) are enumerated
    std::unique_ptr<GenSource> gen;
    if (ctx->cfg.min3) {
        MP_TRY(enumerate_scan(ctx, s, lo, hi, &n));
    } else {
        gen.reset(new GenSource());
        const bool all = lo == 0 && expected_tuples != 0 && expected_tuples == ctx->kmer_total;
        MP_TRY(gen_histograms(ctx, s, lo, hi, scan_pass_mask(ctx->cfg.k), all, gen.get(), &n));
    }
    hipEventRecord(ctx->ev[1], ctx->stream);
    if (expected_tuples && n != expected_tuples)
        return ctx->fail(MP_ECOUNT, "scan %d enumerated %llu tuples, the histogram sized %llu", s,
                         (unsigned long long)n, (unsigned long long)expected_tuples);
    int passes = 0;
    MP_TRY(radix_sort_tuples(ctx, n, &passes, gen.get()));
    hipEventRecord(ctx->ev[7], ctx->stream);
    MP_TRY(filter_union_scan(ctx, n, seg_bounds, T, &st));
    st.tuples = n;
    st.sort_passes = passes;
    st.ms_sort_pass_avg = ctx->sort_pass_avg;
    st.ms_sort_hist = ctx->sort_hist_ms;
    st.ms_sort_local = ctx->local_ms;
    st.sort_long_groups = ctx->local_long;
    st.sort_local_moved = ctx->local_moved;
    st.ms_sort_gen = gen ? ctx->sort_gen_ms : 0.f;
    st.ms_sort_count = ctx->sort_count_ms;
    st.sort_graph = ctx->sort_graph_last;
    hipEventElapsedTime(&st.ms_enumerate, ctx->ev[0], ctx->ev[1]);
    hipEventElapsedTime(&st.ms_sort, ctx->ev[1], ctx->ev[7]);
    if (stats) *stats = st;
    ctx->last_n = n;
    ctx->have_scan = true;
    ctx->scans_done++;
    return 0;
}

int mp_finish(mp_ctx *ctx, uint64_t *W, mp_finish_stats *stats)
{
    if (!ctx || !W) return MP_EINVAL;
    if (ctx->scans_done == 0) return ctx->fail(MP_ESTATE, "mp_finish before any scan");
    MP_HIP(hipSetDevice(ctx->device));
    mp_finish_stats st;
    memset(&st, 0, sizeof(st));
    MP_TRY(finish_components(ctx, W, &st));
    if (stats) *stats = st;
    ctx->finished = true;
    return 0;
}

int mp_get_parent(mp_ctx *ctx, uint64_t *parent)
{
    if (!ctx || !parent) return MP_EINVAL;
    if (ctx->N == 0) return ctx->fail(MP_ESTATE, "no reads loaded");
    MP_HIP(hipSetDevice(ctx->device));
    std::vector<uint32_t> p(ctx->N);
    MP_HIP(hipMemcpyAsync(p.data(), ctx->parent.p, ctx->N * 4, hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    for (uint64_t i = 0; i < ctx->N; i++) parent[i] = p[i];
    return 0;
}

int mp_get_bad(mp_ctx *ctx, uint8_t *bad)
{
    if (!ctx || !bad) return MP_EINVAL;
    if (ctx->N == 0) return ctx->fail(MP_ESTATE, "no reads loaded");
    MP_HIP(hipSetDevice(ctx->device));
    MP_HIP(hipMemcpyAsync(bad, ctx->bad.p, ctx->N, hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int mp_dump_tuples(mp_ctx *ctx, uint8_t *dst, uint64_t cap, uint64_t *n)
{
    if (!ctx || !n) return MP_EINVAL;
    if (!ctx->have_scan) return ctx->fail(MP_ESTATE, "no scan to dump");
    *n = ctx->last_n;
    if (!dst) return 0;
    if (cap < ctx->last_n) return ctx->fail(MP_EINVAL, "dump buffer holds %llu tuples, need %llu",
                                            (unsigned long long)cap, (unsigned long long)ctx->last_n);
    MP_HIP(hipSetDevice(ctx->device));
    const uint64_t m = ctx->last_n;
    std::vector<uint64_t> k(m);
    std::vector<uint32_t> l(m), z((m + 31) / 32 + 1);
    std::vector<uint8_t> a(m);
    MP_HIP(hipMemcpyAsync(k.data(), ctx->keys[ctx->cur].p, m * 8, hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipMemcpyAsync(l.data(), ctx->lab[ctx->cur].p, m * 4, hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipMemcpyAsync(a.data(), ctx->adj[ctx->cur].p, m, hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipMemcpyAsync(z.data(), ctx->zbits.p, z.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    // the device sorts by k-mer only (runs in generation order): order each run
    // by adj, stably, for the reference's (k-mer, adj) order; batches of
    // samples also break runs at sample boundaries
    std::vector<uint64_t> cuts;
    if (ctx->batch) cuts = ctx->sample_tuples;
    std::vector<uint64_t> idx(m);
    for (uint64_t i = 0; i < m; i++) idx[i] = i;
    size_t c = 0;
    for (uint64_t i = 0; i < m;) {
        while (c < cuts.size() && cuts[c] <= i) c++;
        const uint64_t lim = c < cuts.size() ? cuts[c] : m;
        uint64_t j = i + 1;
        while (j < lim && k[j] == k[i]) j++;
        if (j - i > 1)
            std::stable_sort(idx.begin() + i, idx.begin() + j, [&](uint64_t x, uint64_t y) { return a[x] < a[y]; });
        i = j;
    }
    for (uint64_t i = 0; i < m; i++) {  // enumerateKmers.c:352-365 layout
        const uint64_t s = idx[i];
        const uint64_t key = ((z[s >> 5] >> (s & 31)) & 1) ? 0 : k[s];  // zeroed by the filter
        uint8_t *t = dst + 14 * i;
        memcpy(t, &key, 8);
        t[8] = a[s];
        t[9] = 0;  // id / UINT32_MAX: ids are 32-bit here
        memcpy(t + 10, &l[s], 4);
    }
    return 0;
}

// ---------------------------------------------------------------- batches of samples
static int batch_after_load(mp_ctx *ctx)
{
    for (int c = 0; c < ctx->C; c++) {
        const uint64_t nr = ctx->chunk_ids[c + 1] - ctx->chunk_ids[c];
        if (nr == 0 || (nr & 1))
            return ctx->fail(MP_EINVAL, "sample %d holds %llu records (need an even, non-zero count)", c,
                             (unsigned long long)nr);
    }
    ctx->batch = true;
    return 0;
}

int mp_batch_load(mp_ctx *ctx, int n, const char *const *fq, const uint64_t *nbytes)
{
    if (!ctx || n <= 0 || !fq || !nbytes) return MP_EINVAL;
    if (n > MP_MAX_SAMPLES) return ctx->fail(MP_EINVAL, "at most %d samples per batch", MP_MAX_SAMPLES);
    MP_HIP(hipSetDevice(ctx->device));
    std::vector<uint64_t> ofs(n + 1, 0);
    for (int i = 0; i < n; i++) ofs[i + 1] = ofs[i] + nbytes[i];
    MP_TRY(ensure(ctx, ctx->fq, ofs[n] + 64));
    for (int i = 0; i < n; i++)
        if (nbytes[i]) MP_TRY(copy_to_device(ctx, (uint8_t *)ctx->fq.p + ofs[i], fq[i], nbytes[i]));
    MP_TRY(fill_u32(ctx, (uint8_t *)ctx->fq.p + ((ofs[n] + 3) & ~3ull), 0, 8));
    ctx->d_fq = (const uint8_t *)ctx->fq.p;
    ctx->fq_owned = true;
    ctx->fq_bytes = ofs[n];
    MP_TRY(load_common(ctx, ofs.data(), nullptr, n));
    return batch_after_load(ctx);
}

int mp_batch_load_device(mp_ctx *ctx, const void *d_fq, uint64_t nbytes, int n, const uint64_t *sample_ofs)
{
    if (!ctx || !d_fq || n <= 0 || !sample_ofs) return MP_EINVAL;
    if (n > MP_MAX_SAMPLES) return ctx->fail(MP_EINVAL, "at most %d samples per batch", MP_MAX_SAMPLES);
    MP_HIP(hipSetDevice(ctx->device));
    if (ctx->fq_owned) {
        MP_HIP(hipStreamSynchronize(ctx->stream));
        release(ctx, ctx->fq);
    }
    ctx->d_fq = (const uint8_t *)d_fq;
    ctx->fq_owned = false;
    ctx->fq_bytes = nbytes;
    MP_TRY(load_common(ctx, sample_ofs, nullptr, n));
    return batch_after_load(ctx);
}

int mp_batch_partition(mp_ctx *ctx, mp_sample_result *res, mp_scan_stats *scan_stats, mp_finish_stats *finish_stats)
{
    if (!ctx) return MP_EINVAL;
    if (!ctx->batch) return ctx->fail(MP_ESTATE, "mp_batch_partition before mp_batch_load");
    MP_HIP(hipSetDevice(ctx->device));
    const uint64_t N = ctx->N;
    const int ns = ctx->C;
    mp_scan_stats st;
    memset(&st, 0, sizeof(st));
    MP_TRY(iota_u32(ctx, (uint32_t *)ctx->parent.p, N));
    MP_TRY(fill_u32(ctx, ctx->bad.p, 0, (N + 3) / 4));
    MP_TRY(fill_u32(ctx, ctx->filtered.p, 0, (N + 3) / 4));
    ctx->scans_done = 0;
    hipEventRecord(ctx->ev[0], ctx->stream);
    uint64_t n = 0;
    std::unique_ptr<GenSource> gen;
    // every sample sorted on its own (segmented passes, each sample's tuples
    // from its own position) unless the passes take look-back offsets
    const char *lbe = getenv("MP_SORT_LOOKBACK");
    const bool segmented = !ctx->cfg.min3 && !(lbe && atoi(lbe) == 1);
    if (ctx->cfg.min3) {  // every sample over its own one-scan bucket range
        MP_TRY(enumerate_scan(ctx, 0, 0, MP_NBUCKETS, &n, true));
    } else {
        gen.reset(new GenSource());
        if (segmented) MP_TRY(batch_counts(ctx));
        MP_TRY(gen_histograms(ctx, 0, 0, MP_NBUCKETS, scan_pass_mask(ctx->cfg.k), true, gen.get(), &n));
    }
    hipEventRecord(ctx->ev[1], ctx->stream);
    std::vector<uint64_t> tpos(ns + 1, 0);
    int passes = 0;
    float gen_ms = 0.f;
    if (segmented && gen->gcnt && gen->gcnt_seg && gen->seg_tuples.size() == (size_t)ns) {
        for (int i = 0; i < ns; i++) {  // a plain run's range plan is degenerate without k-mers (mp_plan_ranges)
            if (gen->seg_tuples[i] == 0) return ctx->fail(MP_EPLAN, "sample %d has no k-mers to partition", i);
            tpos[i + 1] = tpos[i] + gen->seg_tuples[i];
        }
        SortSegments sg;
        sg.ns = ns;
        sg.tpos = tpos.data();
        sg.wpos = ctx->seg_wpos.data();
        MP_TRY(radix_sort_tuples(ctx, n, &passes, gen.get(), &sg));
        gen_ms = ctx->sort_gen_ms;
    } else {
        // sort by (sample, k-mer, adj): the sample of a label is its chunk
        uint64_t *d_first = nullptr;
        MP_HIP(hipMallocAsync((void **)&d_first, sizeof(uint64_t) * (ns + 1), ctx->stream));
        MP_HIP(hipMemcpyAsync(d_first, ctx->chunk_ids.data(), sizeof(uint64_t) * (ns + 1), hipMemcpyHostToDevice,
                              ctx->stream));
        int rc = radix_sort_tuples(ctx, n, &passes, gen.get());
        gen_ms = gen ? ctx->sort_gen_ms : 0.f;
        if (!rc) {  // then the sample of the label as the most significant digit
            int sp = 0;
            rc = radix_sort_generic(ctx, ctx->keys, ctx->lab, ctx->adj, &ctx->cur, n, 1u << 9, &sp, d_first, ns,
                                    tpos.data());
            passes += sp;
        }
        MP_HIP(hipFreeAsync(d_first, ctx->stream));
        MP_TRY(rc);
    }
    if (n <= 1) {  // the sort returned early: one tuple or none, all in the sample that owns it
        uint32_t l0 = 0;
        if (n == 1) {
            MP_HIP(hipMemcpyAsync(&l0, ctx->lab[ctx->cur].p, 4, hipMemcpyDeviceToHost, ctx->stream));
            MP_HIP(hipStreamSynchronize(ctx->stream));
        }
        for (int i = 0; i <= ns; i++) tpos[i] = (n == 1 && ctx->chunk_ids[i] > l0) ? 1 : 0;
        tpos[ns] = n;
    }
    for (int i = 0; i < ns; i++)  // a plain run's range plan is degenerate without k-mers (mp_plan_ranges)
        if (tpos[i + 1] == tpos[i]) return ctx->fail(MP_EPLAN, "sample %d has no k-mers to partition", i);
    hipEventRecord(ctx->ev[7], ctx->stream);
    // one filter/union segment per sample, each its own quirk group
    std::vector<uint64_t> grp(2 * (size_t)ns);
    for (int i = 0; i < ns; i++) {
        grp[2 * i] = tpos[i];
        grp[2 * i + 1] = tpos[i + 1];
    }
    MP_TRY(filter_union_groups(ctx, n, tpos.data(), ns, grp.data(), &st));
    st.tuples = n;
    st.sort_passes = passes;
    st.ms_sort_pass_avg = ctx->sort_pass_avg;
    st.ms_sort_hist = ctx->sort_hist_ms;
    st.ms_sort_local = ctx->local_ms;
    st.sort_long_groups = ctx->local_long;
    st.sort_local_moved = ctx->local_moved;
    st.ms_sort_gen = gen_ms;
    st.ms_sort_count = ctx->sort_count_ms;
    st.sort_graph = ctx->sort_graph_last;
    hipEventElapsedTime(&st.ms_enumerate, ctx->ev[0], ctx->ev[1]);
    hipEventElapsedTime(&st.ms_sort, ctx->ev[1], ctx->ev[7]);
    ctx->last_n = n;
    ctx->have_scan = true;
    ctx->scans_done = 1;
    ctx->sample_tuples = tpos;
    mp_finish_stats fs;
    memset(&fs, 0, sizeof(fs));
    uint64_t W[MP_NPART + 1];
    MP_TRY(finish_components(ctx, W, &fs));
    ctx->finished = true;
    if (res) {
        for (int i = 0; i < ns; i++) {
            mp_sample_result &r = res[i];
            r.reads = ctx->chunk_ids[i + 1] - ctx->chunk_ids[i];
            r.tuples = tpos[i + 1] - tpos[i];
            r.components = ctx->sample_comp[i];
            r.largest = ctx->sample_largest[i];
            for (int p = 0; p <= MP_NPART; p++) r.W[p] = ctx->sample_W[(size_t)i * (MP_NPART + 1) + p];
        }
    }
    if (scan_stats) *scan_stats = st;
    if (finish_stats) *finish_stats = fs;
    return 0;
}

int mp_batch_get_parent(mp_ctx *ctx, int s, uint32_t *parent)
{
    if (!ctx || !parent) return MP_EINVAL;
    if (!ctx->batch || !ctx->finished) return ctx->fail(MP_ESTATE, "no partitioned batch");
    if (s < 0 || s >= ctx->C) return ctx->fail(MP_EINVAL, "sample %d out of range", s);
    MP_HIP(hipSetDevice(ctx->device));
    const uint64_t a = ctx->chunk_ids[s], b = ctx->chunk_ids[s + 1];
    MP_HIP(hipMemcpyAsync(parent, (const uint32_t *)ctx->parent.p + a, (b - a) * 4, hipMemcpyDeviceToHost,
                          ctx->stream));
    MP_HIP(hipStreamSynchronize(ctx->stream));
    for (uint64_t i = 0; i < b - a; i++) parent[i] -= (uint32_t)a;
    return 0;
}

int mp_batch_write(mp_ctx *ctx, const char *out_dir, uint64_t *bytes_out)
{
    if (!ctx) return MP_EINVAL;
    if (!ctx->batch || !ctx->finished) return ctx->fail(MP_ESTATE, "no partitioned batch");
    MP_HIP(hipSetDevice(ctx->device));
    return write_partitions(ctx, out_dir, 1, nullptr, nullptr, bytes_out);
}

int mp_write_partitions(mp_ctx *ctx, const char *out_prefix, int T, const uint64_t *W, const int64_t *il_start,
                        uint64_t *bytes_out)
{
    if (!ctx || !W || T <= 0) return MP_EINVAL;
    if (!ctx->finished) return ctx->fail(MP_ESTATE, "mp_write_partitions before mp_finish");
    if (ctx->C % T) return ctx->fail(MP_EINVAL, "T must divide the chunk count");
    MP_HIP(hipSetDevice(ctx->device));
    return write_partitions(ctx, out_prefix, T, W, il_start, bytes_out);
}

}  // extern "C"
