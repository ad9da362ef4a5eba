// Gather of byte ranges laid end to end: item i of n is a range of a source
// buffer that lands at [dst_off[i], dst_off[i+1]) of a destination space
// (dst_off non-decreasing).  Two users: the reads' sequence bytes, packed into
// the 2-bit and 'N' streams as they are gathered, and the writer's records,
// copied in file order.
//
// A workgroup owns 16 KiB of the destination space, so no two workgroups
// write the same byte or word.  The items that overlap its range (their first
// is found by gather_first_kernel) are staged in LDS as (start, source delta),
// and each lane takes 16-byte destination units: a binary search in LDS finds
// the unit's item; the unit's bytes of that item are read as two aligned
// 16-byte loads and shifted into place, and a unit across an item end reads
// the next item's bytes the same way with its loads issued beside the first
// (a byte-by-byte tail only for units over three or more items).  The
// earlier form, one wave per item, was bound by each wave's dependent loads
// (1.6 TB/s on 150-315-byte items).
#pragma once
#include "mp_internal.h"

namespace mp {

constexpr int kGLanes = 256;
constexpr uint32_t kGUnits = 1024;                 // 16-byte units per workgroup
constexpr uint64_t kGBlock = 16ull * kGUnits;      // destination bytes per workgroup
constexpr uint32_t kGItems = 1024;                 // items staged in LDS (more: searched in memory)

// first[b] = the item holding destination byte b * kGBlock, for the nb blocks
// that hold data (b * kGBlock < dst_off[n]): the largest i with dst_off[i] <= b * kGBlock.
static __global__ void gather_first_kernel(const uint64_t *__restrict__ dst_off, uint64_t n, uint64_t nb,
                                    uint64_t *__restrict__ first)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t lo = dst_off[i], hi = dst_off[i + 1];
        for (uint64_t b = (lo + kGBlock - 1) / kGBlock; b < nb && b * kGBlock < hi; b++) first[b] = i;
    }
}

// 16 source bytes at src[s .. s+16) (absolute offset s into a buffer of nbytes) as 4 words.
// Two aligned 16-byte loads when they stay inside the buffer, else bytes.
__device__ __forceinline__ void load16(const uint8_t *__restrict__ src, uint64_t nbytes, uint64_t s, uint32_t out[4])
{
    const uintptr_t a = (uintptr_t)(src + s);
    const uintptr_t al = a & ~(uintptr_t)15;
    if (al >= (uintptr_t)src && al + 32 <= (uintptr_t)(src + nbytes)) {
        const uint4 p = *(const uint4 *)al, q = *(const uint4 *)(al + 16);
        uint32_t w[8] = {p.x, p.y, p.z, p.w, q.x, q.y, q.z, q.w};
        const uint32_t o = (uint32_t)(a & 15);
        if (o & 8) {
#pragma unroll
            for (int k = 0; k < 6; k++) w[k] = w[k + 2];
        }
        if (o & 4) {
#pragma unroll
            for (int k = 0; k < 5; k++) w[k] = w[k + 1];
        }
        const uint32_t sh = 8 * (o & 3);
#pragma unroll
        for (int k = 0; k < 4; k++) out[k] = (uint32_t)((((uint64_t)w[k + 1] << 32) | w[k]) >> sh);
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint32_t v = 0;
            for (int e = 0; e < 4; e++)
                if (s + 4 * k + e < nbytes) v |= (uint32_t)src[s + 4 * k + e] << (8 * e);
            out[k] = v;
        }
    }
}

// bytes [0, e) of a 16-byte unit within its word k (e <= 16), as a mask
__device__ __forceinline__ uint32_t prefix_mask(uint32_t e, int k)
{
    const uint32_t lo = 4u * k;
    return e >= lo + 4 ? 0xFFFFFFFFu : e <= lo ? 0u : (1u << (8 * (e - lo))) - 1;
}

// Sink of the sequence gather: 16 bases at base index x (a multiple of 16)
// become one 32-bit half of a 2-bit word (first base most significant) and
// one 16-bit quarter of an 'N' word; bytes past the stream (valid = 0) are
// zero codes without an 'N'.
struct PackSink {
    uint32_t *packed;  // the 2-bit stream as 32-bit halves
    uint16_t *nmask;   // the 'N' stream as 16-bit quarters
    __device__ void operator()(uint64_t x, const uint32_t b[4], uint32_t valid) const
    {
        uint32_t code = 0, nb = 0;
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const uint32_t c = (b[e >> 2] >> (8 * (e & 3))) & 0xFF;
            if (e < (int)valid) {
                code |= ((0x2310u >> (((c & 6) >> 1) * 4)) & 3) << (30 - 2 * e);
                if (c == 'N') nb |= 1u << (15 - e);
            }
        }
        const uint64_t u = x >> 4;
        packed[2 * (u >> 1) + 1 - (u & 1)] = code;  // bases 0-15 of a word are its high half
        nmask[4 * (u >> 2) + 3 - (u & 3)] = (uint16_t)nb;
    }
};

// Sink of the record gather: plain bytes.
struct CopySink {
    uint8_t *out;
    __device__ void operator()(uint64_t x, const uint32_t b[4], uint32_t valid) const
    {
        if (valid == 16) {
            *(uint4 *)(out + x) = make_uint4(b[0], b[1], b[2], b[3]);
        } else {
#pragma unroll
            for (uint32_t e = 0; e < 16; e++)
                if (e < valid) out[x + e] = (uint8_t)(b[e >> 2] >> (8 * (e & 3)));
        }
    }
};

// One workgroup per 16 KiB block of the destination space [0, span): bytes
// below total = dst_off[n] are gathered, bytes from total to span go to the
// sink as invalid (the packed streams' zero tail).  src_of(i) is item i's
// source offset.  nb = blocks that hold data; later blocks only pad.
template <typename SrcOf, typename Sink>
__global__ __launch_bounds__(kGLanes) void gather_kernel(const uint8_t *__restrict__ src, uint64_t nbytes,
                                                         const uint64_t *__restrict__ dst_off, uint64_t n,
                                                         const uint64_t *__restrict__ first, uint64_t nb,
                                                         uint64_t total, uint64_t span, SrcOf src_of, Sink sink)
{
    __shared__ uint32_t s_beg[kGItems + 1];  // item starts relative to the block (the first clamped to 0)
    __shared__ uint64_t s_delta[kGItems];    // source offset minus destination offset
    const uint32_t tid = threadIdx.x;
    const uint64_t b = blockIdx.x, x0 = b * kGBlock;
    const uint32_t lim_out = (uint32_t)min(kGBlock, span - x0);
    const uint32_t lim = x0 < total ? (uint32_t)min(kGBlock, total - x0) : 0u;
    uint64_t i0 = 0;
    uint32_t m = 0;  // items overlapping the block's data
    if (lim) {
        i0 = first[b];
        const uint64_t i1 = b + 1 < nb ? first[b + 1] : n - 1;
        m = (uint32_t)(i1 - i0 + 1);
    }
    const bool staged = m <= kGItems;
    if (staged) {
        for (uint32_t j = tid; j < m; j += kGLanes) {
            const uint64_t d = dst_off[i0 + j];
            s_beg[j] = d > x0 ? (uint32_t)(d - x0) : 0u;
            s_delta[j] = src_of(i0 + j) - d;
        }
        if (tid == 0) s_beg[m] = lim;
        __syncthreads();
    }
    auto beg = [&](uint32_t j) -> uint32_t {
        if (staged) return s_beg[j];
        if (j == m) return lim;
        const uint64_t d = dst_off[i0 + j];
        return d > x0 ? (uint32_t)(d - x0) : 0u;
    };
    auto delta = [&](uint32_t j) -> uint64_t {
        if (staged) return s_delta[j];
        return src_of(i0 + j) - dst_off[i0 + j];
    };
    for (uint32_t y = 16 * tid; y < lim_out; y += 16 * kGLanes) {
        uint32_t w[4] = {0, 0, 0, 0};
        const uint32_t valid = y < lim ? min(16u, lim - y) : 0u;
        if (valid) {
            // item of byte y: the last j < m with beg(j) <= y
            uint32_t lo = 0, hi = m - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (beg(mid) <= y) lo = mid; else hi = mid - 1;
            }
            // the unit's bytes come from item j, then (past an item end) from the
            // next non-empty item, both read as 16 aligned-load bytes at once;
            // a unit over three or more items takes its remaining bytes singly
            uint32_t j = lo;
            const uint32_t eA = min(beg(j + 1) - y, valid);
            load16(src, nbytes, x0 + y + delta(j), w);
            if (eA < valid) {
                uint32_t jb = j + 1;
                while (beg(jb + 1) <= y + eA) jb++;  // zero-length items
                uint32_t t[4];
                load16(src, nbytes, x0 + y + delta(jb), t);
                const uint32_t eB = min(beg(jb + 1) - y, valid);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t a = prefix_mask(eA, k), bm = prefix_mask(eB, k) & ~a;
                    w[k] = (w[k] & a) | (t[k] & bm);
                }
                j = jb;
#pragma unroll
                for (uint32_t e = 0; e < 16; e++) {
                    if (e >= eB && e < valid) {
                        while (beg(j + 1) <= y + e) j++;
                        w[e >> 2] |= (uint32_t)src[x0 + y + e + delta(j)] << (8 * (e & 3));
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++) w[k] &= prefix_mask(valid, k);
            }
        }
        sink(x0 + y, w, valid);
    }
}

// Gathers items [0, n) (destination offsets dst_off[0..n], source offsets
// src_of(i)) over the destination space [0, span), span >= dst_off[n]; the
// caller passes total = dst_off[n] (read back already).  first: scratch of
// ceil(total / 16 KiB) words.
template <typename SrcOf, typename Sink>
static int launch_gather(mp_ctx *ctx, const uint8_t *src, uint64_t nbytes, const uint64_t *dst_off, uint64_t n,
                         uint64_t total, uint64_t span, uint64_t *first, SrcOf src_of, Sink sink)
{
    const uint64_t nb = (total + kGBlock - 1) / kGBlock, nspan = (span + kGBlock - 1) / kGBlock;
    if (nspan == 0) return 0;
    if (nb && n == 0) return ctx->fail(MP_EINVAL, "gather: %llu bytes from no items", (unsigned long long)total);
    if (!grid_fits(nspan, kGLanes)) return ctx->fail(MP_EINVAL, "gather: %llu bytes in one launch", (unsigned long long)span);
    if (nb) {
        hipLaunchKernelGGL(gather_first_kernel, dim3(stride_grid(n)), dim3(256), 0, ctx->stream, dst_off, n, nb, first);
        if (hipGetLastError() != hipSuccess) return ctx->fail(MP_EDEVICE, "gather_first_kernel launch");
    }
    hipLaunchKernelGGL((gather_kernel<SrcOf, Sink>), dim3((unsigned)nspan), dim3(kGLanes), 0, ctx->stream, src, nbytes,
                       dst_off, n, (const uint64_t *)first, nb, total, span, src_of, sink);
    if (hipGetLastError() != hipSuccess) return ctx->fail(MP_EDEVICE, "gather_kernel launch");
    return 0;
}

}  // namespace mp
