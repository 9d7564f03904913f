// draw_lines on gfx950: segment checkpoints, tile-binned LDS accumulation,
// pixel lists, the deterministic (reference-order) mode, accumulate_lines,
// plans and the adjoint.
//
// Tile path (draw_lines, plans; F frames of H x W, N lines):
//
//  1. k_setup_lines   one lane per line: the line's constants (LineSetup) and
//                     its segment bound; the bound total (the slot count;
//                     an eager draw reads it first only when the previous
//                     call's total is no usable capacity).
//  2. k_walk_lines    one lane per line: replays the reference's walk once,
//                     cutting it into segments of <= SEG long-axis columns
//                     inside the frame.  A segment stores the exact walk
//                     state at its first column (the checkpoint) and the
//                     short-axis pixel range its fills can reach; the walk
//                     counts (segment, tile) items per tile sub-list, one
//                     atomic per run of items into one sub-list.
//  3. k_scan          item offsets per tile sub-list (one host read: the
//                     item total, flags and slot count; plans read nothing).
//  4. k_bin_slots     one lane per segment slot: the per-tile item lists.
//  5. k_tiles         one workgroup per TILE x TILE pixel tile: one wave
//                     resumes each item's walk from its checkpoint (phase 1,
//                     column records into LDS), then all waves fill the
//                     columns (phase 2), accumulating into an LDS tile with
//                     LDS atomics (fixed point for float sums); the tile is
//                     flushed to HBM once with a coalesced read-modify-write.
//                     HBM traffic of the image: one read and one write per
//                     touched tile.
//
// Ordered path (pixel lists, the deterministic mode, accumulate_lines with
// different overlaps): k_segments walks twice (count, then write) so that a
// line's segments are consecutive and lines keep their input order, and
// k_tiles_ordered applies each pixel's terms in ascending line order, the
// single-threaded reference's order.  The adjoint (k_adjoint_tiles,
// k_line_grad) runs on the tile path's segments and lists.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>

#include "cbx_core.hpp"
#include "cbx_host.hpp"

namespace cbx {

template <typename T>
struct Segment {
    int32_t line;
    int32_t c[2];   // walk position at the segment's first column
    int32_t ncols;  // long-axis columns covered from c[axis] in the walk direction
    T Te, Ne;       // walk errors at that position
    int32_t lo, hi; // short-axis pixel range the fills can reach, clipped to the frame
};

struct Geometry {
    int64_t n_frames;
    int32_t H, W;
    int32_t tile, seg;
    int32_t ntx, nty;
    int32_t skip_neg = 0;   // 1: a line with a negative frame draws nothing (accumulate_lines term batches)
};

// ---------------------------------------------------------------------------
// segment pre-passes
// ---------------------------------------------------------------------------
// Line constants and frame check of line l; false if the line draws nothing.
// kid: the line's profile, kernels[l] when per-line profiles are given.
template <typename T, typename I>
__device__ __forceinline__ bool setup_line(const T *__restrict__ lines, const I *__restrict__ idxs,
                                           const uint8_t *__restrict__ kernels, int kernel, int64_t l,
                                           const Geometry &g, LineSetup<T> &L, int *__restrict__ flags)
{
    int rc = line_setup(lines + 5 * l, L);
    const int kid = kernels ? (int)kernels[l] : kernel;
    L.kid = (int8_t)kid;
    if (kid < K_BIWEIGHT || kid > K_TRIANGULAR) {
        atomicOr(flags, FLAG_KERNEL);
        L.valid = 0;
    }
    int64_t f = idxs ? (int64_t)idxs[l] : ((g.n_frames == 1) ? 0 : l);
    if (f < 0 && g.skip_neg) {
        L.valid = 0;
        f = 0;
    } else if (f < 0 || f >= g.n_frames) {
        atomicOr(flags, FLAG_INDEX);
        L.valid = 0;
    }
    if (rc) {
        atomicOr(flags, rc);
        L.valid = 0;
    }
    L.frame = (int32_t)f;
    return L.valid;
}

// Most segments a line's walk can make: one per SEG-column block of the
// long-axis range [m0, m1] inside the frame (the walk is monotone along the
// long axis and never passes m1).
template <typename T>
__device__ __forceinline__ int segment_bound(const LineSetup<T> &L, const Geometry &g)
{
    if (!L.valid) return 0;
    const int a = L.axis, da = (a == 0) ? g.W : g.H;
    const int c0 = pick(L.m0, a), c1 = pick(L.m1, a);
    const int lo = max(min(c0, c1), 0), hi = min(max(c0, c1), da - 1);
    return (lo > hi) ? 0 : hi / g.seg - lo / g.seg + 1;
}

// Replays line l's walk (bresenham.hpp:352-383) and cuts it into segments of
// <= SEG long-axis columns inside the frame.  For each segment whose fills
// can reach the frame, on_seg(cur, tl) gets the checkpoint (with the reach
// range lo..hi set) and the tile column tl along the long axis.
template <typename T, typename OnSeg>
__device__ __forceinline__ void walk_segments(const LineSetup<T> &L, int64_t l, const Geometry &g,
                                              int *__restrict__ flags, OnSeg on_seg)
{
    const int a = L.axis, b = 1 - a;
    const bool ax0 = a == 0;
    const int da = (a == 0) ? g.W : g.H, db = (a == 0) ? g.H : g.W;
    const int sa = pick(L.st, a);
    const double ta = (double)m_abs(pick(L.t, a));
    const double wl = (double)L.w * (double)L.len;

    WalkState<T> s;
    s.c[0] = L.m0[0];
    s.c[1] = L.m0[1];
    s.Te = L.e0[0];
    s.Ne = L.e0[1];

    int b_lo = 0, b_hi = 0, tl = 0, nem = 0, smin = 0, smax = 0;
    T maxN = T(0);   // the largest |N| at an emitted column (exact in T)
    Segment<T> cur;

    auto close = [&]() {
        if (nem == 0) return;
        // fills stop once |N| / len >= w; |N| grows by |t_a| per short step
        int R = (int)((wl + (double)maxN) / ta) + 2;
        int lo = max(smin - R, 0), hi = min(smax + R, db - 1);
        if (lo > hi) return;
        cur.lo = lo;
        cur.hi = hi;
        on_seg(cur, tl);
    };

    // every step moves the walk; a walk takes at most one short-axis-only
    // step per long-axis step, so this budget is never reached by a valid line
    const int64_t budget = 4 * ((int64_t)abs(pick(L.m1, a) - pick(L.m0, a)) + abs(pick(L.m1, b) - pick(L.m0, b))) + 64;
    // Loop structure: an outer loop over segments and an inner loop over the
    // steps of one SEG-column block, so that the lanes of a wave close their
    // segments together (a close inside the step loop would run, with a few
    // lanes active, on almost every step).  The steps and decisions are those
    // of one plain loop over the walk.
    wait_vm_loads();
    int64_t it = 0;
    bool stop = false;
    // columns before the frame (the walk is monotone along the long axis)
    for (;; it++) {
        if (it == budget) {
            atomicOr(flags, FLAG_STUCK);
            stop = true;
            break;
        }
        if (walk_done(L, s)) {
            stop = true;
            break;
        }
        const int c = pick(s.c, a);
        if ((sa > 0) ? (c >= da) : (c < 0)) { // no in-frame column is left
            stop = true;
            break;
        }
        if (c >= 0 && c < da) break; // the first segment starts here
        bool nx, ny;
        walk_next(L, s.Ne, nx, ny);
        walk_step(L, s, nx, ny);
    }
    while (!stop) {
        // open a segment at the current (in-frame) column
        const int c0 = pick(s.c, a);
        b_lo = (c0 / g.seg) * g.seg;
        b_hi = min(b_lo + g.seg, da);
        tl = c0 / g.tile;
        nem = 0;
        maxN = T(0);
        smin = INT_MAX;
        smax = INT_MIN;
        cur.line = (int32_t)l;
        cur.c[0] = s.c[0];
        cur.c[1] = s.c[1];
        cur.Te = s.Te;
        cur.Ne = s.Ne;
        cur.ncols = (sa > 0) ? (b_hi - c0) : (c0 - b_lo + 1);
        // One exit test per step: the long-axis coordinate moves by 0 or sa
        // per step and starts inside the block [b_lo, b_hi), which lies
        // inside the frame, so the block (and the frame) is left exactly at
        // c_end.  Steps are counted per block in 32 bits (at most two per
        // column for a valid walk: kmax is never reached) and added to the
        // walk's budget after it.  Why the block ended is sorted out after
        // the loop, in the order the checks have in a plain loop: the step
        // budget, the end of the walk, the frame edge, else the next block.
        // Branch-free updates keep the step short.
        const int c_end = (sa > 0) ? b_hi : b_lo - 1;
        const int kmax = 4 * g.seg + 8;
        int k = 0;
        for (;; k++) {
            if ((k == kmax) | (pick(s.c, a) == c_end) | walk_done(L, s)) break;
            bool nx, ny;
            walk_next(L, s.Ne, nx, ny);
            // (lane masks combined with the loop-invariant axis mask; a NaN
            // |N| leaves maxN, as fmax would)
            const bool em = (nx && ax0) || (ny && !ax0);
            const int sb = pick(s.c, b);
            const T an = m_abs(s.Ne);
            nem += em ? 1 : 0;
            smin = em ? min(smin, sb) : smin;
            smax = em ? max(smax, sb) : smax;
            maxN = (em && an > maxN) ? an : maxN;
            walk_step(L, s, nx, ny);
        }
        it += k;
        if (k == kmax || it >= budget) {
            atomicOr(flags, FLAG_STUCK);
            stop = true;
        } else if (walk_done(L, s)) {
            stop = true;
        } else {
            const int c = pick(s.c, a);
            if ((sa > 0) ? (c >= da) : (c < 0)) stop = true;
        }
        close();
    }
}

__device__ __forceinline__ int64_t tile_index(const Geometry &g, int frame, int axis, int tl, int tb)
{
    const int tx = (axis == 0) ? tl : tb, ty = (axis == 0) ? tb : tl;
    return ((int64_t)frame * g.nty + ty) * g.ntx + tx;
}

// Ordered plan (pixel lists, deterministic images, adjoint): segments of a
// line are consecutive and lines are in input order.  Pass 1 (WRITE = false)
// counts segments per line and items per tile; pass 2 writes them at the
// scanned offsets.
template <typename T, typename I, bool WRITE, bool BIN>
__global__ void __launch_bounds__(256)
k_segments(const T *__restrict__ lines, const I *__restrict__ idxs, int kernel, int64_t n_lines, Geometry g,
           LineSetup<T> *__restrict__ setup, int *__restrict__ seg_cnt, int *__restrict__ tile_cnt,
           const int *__restrict__ seg_off, Segment<T> *__restrict__ segs, const int *__restrict__ tile_off,
           int *__restrict__ tile_cur, int *__restrict__ items, int *__restrict__ flags)
{
    int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n_lines) return;
    LineSetup<T> L;
    if (!WRITE) {
        const bool ok = setup_line(lines, idxs, (const uint8_t *)nullptr, kernel, l, g, L, flags);
        setup[l] = L;
        if (!ok) {
            seg_cnt[l] = 0;
            return;
        }
    } else {
        L = setup[l];
        if (!L.valid) return;
    }
    const int base = WRITE ? seg_off[l] : 0;
    int nseg = 0;
    walk_segments(L, l, g, flags, [&](const Segment<T> &cur, int tl) {
        const int sid = base + nseg;
        if (WRITE) segs[sid] = cur;
        if (BIN) {
            for (int tb = cur.lo / g.tile; tb <= cur.hi / g.tile; tb++) {
                const int64_t tid = tile_index(g, L.frame, L.axis, tl, tb);
                if (!WRITE) atomicAdd(&tile_cnt[tid], 1);
                else items[tile_off[tid] + atomicAdd(&tile_cur[tid], 1)] = sid;
            }
        }
        nseg++;
    });
    if (!WRITE) seg_cnt[l] = nseg;
}

// Unordered plan (the tile path, whose result does not depend on segment
// order): one walk per line.
//   k_setup_lines   line constants, flags, and the total of the per-line
//                   segment bounds (counters[0]);
//   k_walk_lines    orders the workgroup's lines by segment bound, takes
//                   slots for their bounds with one atomic (counters[1]),
//                   walks, writes the segments into them (unused slots get
//                   ncols = 0) and counts tile items (atomics the walk does
//                   not wait for);
//   k_bin_slots     one lane per slot: writes the per-tile item lists, one
//                   cursor atomic per (wave, tile).
template <typename T, typename I>
__global__ void __launch_bounds__(256)
k_setup_lines(const T *__restrict__ lines, const I *__restrict__ idxs, const uint8_t *__restrict__ kernels,
              int kernel, int64_t n_lines, Geometry g, LineSetup<T> *__restrict__ setup,
              unsigned long long *__restrict__ counters, int *__restrict__ flags)
{
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int ub = 0;
    if (l < n_lines) {
        LineSetup<T> L;
        setup_line(lines, idxs, kernels, kernel, l, g, L, flags);
        setup[l] = L;
        ub = segment_bound(L, g);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ub += __shfl_xor(ub, o, 64);
    // one global atomic per workgroup (same-address atomics serialise)
    __shared__ unsigned long long bsum;
    if (threadIdx.x == 0) bsum = 0ull;
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && ub) atomicAdd(&bsum, (unsigned long long)ub);
    __syncthreads();
    if (threadIdx.x == 0 && bsum) atomicAdd(&counters[0], bsum);
}

// The unordered plan's item list of tile t is split into LIST_SUBS sub-lists
// by the binning wave (slot / 64 mod LIST_SUBS), so that the cursor atomics
// of a dense tile spread over LIST_SUBS counters; the sub-lists are
// contiguous: tile t's items are [tile_off[t * LIST_SUBS],
// tile_off[(t + 1) * LIST_SUBS]).
constexpr int LIST_SUBS = 8;

__device__ __forceinline__ int64_t list_index(int64_t tile, int64_t slot)
{
    return tile * LIST_SUBS + ((slot >> 6) & (LIST_SUBS - 1));
}

// Records per thread in the walk's LDS ring: 2 (16 KB per 256 threads; a
// ring of 4 gave no faster walk on 128k lines and fewer resident workgroups
// on 972k lines).
constexpr int WALK_RING = 2;

template <typename T>
__global__ void __launch_bounds__(256)
k_walk_lines(int64_t n_lines, Geometry g, const LineSetup<T> *__restrict__ setup, Segment<T> *__restrict__ segs,
             int64_t slot_cap, unsigned long long *__restrict__ counters, int *__restrict__ tile_cnt,
             int *__restrict__ flags, int64_t *__restrict__ line_base)
{
    const int64_t l0 = (int64_t)blockIdx.x * blockDim.x;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6, n_waves = (blockDim.x + 63) >> 6;
    LineSetup<T> L;
    int ub = 0;
    int64_t l = l0 + threadIdx.x;
    if (l < n_lines) {
        L = setup[l];
        ub = segment_bound(L, g);
    }
    {
        // A wave steps until its longest walk ends.  The workgroup's lines
        // are ordered by segment bound (min(bound, 63); a stable LSD radix
        // sort of one bit per pass, ballot splits), so that each wave walks
        // lines of similar length; thread t then walks the t-th line.
        __shared__ int s_ub[256], s_id[256], s_zero[4];
        int id = threadIdx.x;
        const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
        for (int bit = 0; bit < 6; bit++) {
            const bool one = (min(ub, 63) >> bit) & 1;
            const uint64_t zm = __ballot(!one);
            const int zb = __popcll(zm & below);
            if (lane == 0) s_zero[wave] = __popcll(zm);
            __syncthreads();
            int zt = 0, zw = 0;
            for (int w = 0; w < n_waves; w++) {
                const int z = s_zero[w];
                zw += (w < wave) ? z : 0;
                zt += z;
            }
            // zeros keep their order in front, ones after them
            const int pos = one ? zt + (wave * 64 - zw) + (lane - zb) : zw + zb;
            s_ub[pos] = ub;
            s_id[pos] = id;
            __syncthreads();
            ub = s_ub[threadIdx.x];
            id = s_id[threadIdx.x];
        }
        l = l0 + id;
        if (ub > 0) L = setup[l];
    }
    // wave prefix of the bounds, one slot reservation per wave
    int pre = ub;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(pre, d, 64);
        if (lane >= d) pre += y;
    }
    const int total = __shfl(pre, 63, 64);
    pre -= ub;
    // slots for the workgroup's waves with one global atomic (same-address
    // atomics serialise), handed out in wave order
    __shared__ unsigned long long wsum[4];
    if (lane == 0) wsum[wave] = (unsigned long long)total;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long bt = 0;
        for (int w = 0; w < n_waves; w++) bt += wsum[w];
        unsigned long long b = bt ? atomicAdd(&counters[1], bt) : 0ull;
        if (bt && (int64_t)(b + bt) > slot_cap) atomicOr(flags, FLAG_OVERFLOW);
        for (int w = 0; w < n_waves; w++) {
            const unsigned long long t = wsum[w];
            wsum[w] = b;
            b += t;
        }
    }
    __syncthreads();
    const unsigned long long wbase = wsum[wave];
    // first slot of each line (the adjoint's per-line reduction)
    if (line_base && l < n_lines) line_base[l] = (int64_t)wbase + pre;
    if (ub == 0 || (int64_t)(wbase + total) > slot_cap) return;
    const int64_t base = (int64_t)wbase + pre;
    int nseg = 0;
    // Segments go through a per-thread ring of WALK_RING records in LDS and
    // are written to memory WALK_RING at a time, all loads before all
    // stores: a store's data registers cannot be reused before the store
    // completes, so writing each segment from registers at its close would
    // stall every close on the previous store.  The tile counts are atomics
    // whose results the walk does not use (k_bin_slots places the items);
    // their rate bounds the walk, so each lane adds up runs of items into
    // the same list (consecutive segments of a line mostly share their
    // tiles) in two pending counters and issues one atomic per run.
    // (ring words [slot][word][thread]: a record moves as 32-bit words with
    // constant indices, so it stays in registers)
    constexpr int NW = (int)(sizeof(Segment<T>) / 4);
    static_assert(sizeof(Segment<T>) % 8 == 0, "records move as 64-bit pairs");
    __shared__ uint32_t ring[WALK_RING][NW][256];
    const int t = threadIdx.x;
    auto flush = [&](int first, int cnt) {   // records first .. first + cnt - 1, ring slots 0 .. cnt - 1
        uint32_t w[WALK_RING][NW];
#pragma unroll
        for (int j = 0; j < WALK_RING; j++)
#pragma unroll
            for (int k = 0; k < NW; k++)
                if (j < cnt) w[j][k] = ring[j][k][t];
#pragma unroll
        for (int j = 0; j < WALK_RING; j++) {
            if (j < cnt && first + j < ub) {
                uint2 *dst = reinterpret_cast<uint2 *>(segs + base + first + j);
#pragma unroll
                for (int k = 0; k < NW / 2; k++) dst[k] = make_uint2(w[j][2 * k], w[j][2 * k + 1]);
            }
        }
    };
    int pend_a = -1, pend_b = -1, n_a = 0, n_b = 0;   // pending list counts (list index, items)
    walk_segments(L, l, g, flags, [&](const Segment<T> &cur, int tl) {
        uint32_t cw[NW];
        __builtin_memcpy(cw, &cur, sizeof(cw));
        const int slot = nseg % WALK_RING;
#pragma unroll
        for (int k = 0; k < NW; k++) ring[slot][k][t] = cw[k];
        const int tb0 = cur.lo / g.tile, tb1 = cur.hi / g.tile;
        for (int tb = tb0; tb <= tb1; tb++) {
            const int li = (int)list_index(tile_index(g, L.frame, L.axis, tl, tb), base + nseg);
            if (li == pend_a) {
                n_a++;
            } else if (li == pend_b) {
                n_b++;
            } else {
                if (n_b) atomicAdd(&tile_cnt[pend_b], n_b);
                pend_b = pend_a;
                n_b = n_a;
                pend_a = li;
                n_a = 1;
            }
        }
        nseg++;
        if (nseg % WALK_RING == 0) flush(nseg - WALK_RING, WALK_RING);
    });
    if (n_a) atomicAdd(&tile_cnt[pend_a], n_a);
    if (n_b) atomicAdd(&tile_cnt[pend_b], n_b);
    if (nseg % WALK_RING) {
        const int done = (nseg / WALK_RING) * WALK_RING;   // records flushed at closes
        flush(done, nseg - done);
    }
    if (nseg > ub) atomicOr(flags, FLAG_STUCK); // cannot happen: the bound holds for every walk
    for (int k = nseg; k < ub; k++) segs[base + k].ncols = 0;
}

// Clears a plan's counters before a draw (a kernel, not a memset: a draw may
// be captured in a HIP graph).
__global__ void k_plan_reset(int64_t n_lists, unsigned long long *__restrict__ counters, int *__restrict__ flags,
                             int *__restrict__ cnt, int *__restrict__ cur, unsigned long long *__restrict__ scan_state,
                             int64_t scan_words)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < scan_words) scan_state[i] = 0ull;
    if (i == 0) {
        counters[0] = 0ull;
        counters[1] = 0ull;
        *flags = 0;
    }
    if (i < n_lists) {
        cnt[i] = 0;
        cur[i] = 0;
    }
}

// Tile t's list: the lanes of a wave (64 consecutive slots) whose segments
// reach one tile take their places with one cursor atomic, in lane order, so
// a list holds runs of consecutive slots (the tile kernels' segment loads
// coalesce) and same-address atomics are few.
template <typename T>
__global__ void __launch_bounds__(256)
k_bin_slots(int64_t n_slots, Geometry g, const LineSetup<T> *__restrict__ setup, const Segment<T> *__restrict__ segs,
            const int *__restrict__ tile_off, int *__restrict__ cur, int2 *__restrict__ pairs,
            const unsigned long long *__restrict__ counters, int64_t n_tiles, int64_t item_cap,
            int *__restrict__ flags, int *__restrict__ slot_tiles)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    // after a slot overflow some reserved slots were never written (stale
    // segments): bin nothing; otherwise the slots taken by the walk, and the
    // item total against the list capacity
    if (*flags & FLAG_OVERFLOW) return;
    if ((int64_t)tile_off[n_tiles * LIST_SUBS] > item_cap) {
        if (i == 0) atomicOr(flags, FLAG_OVERFLOW);
        return;
    }
    int np = 0, axis = 0, frame = 0, tl = 0, tb0 = 0;
    int2 item = make_int2(0, 0);
    if (i < n_slots && i < (int64_t)counters[1]) {
        const Segment<T> sg = segs[i];
        if (sg.ncols > 0) {
            axis = setup[sg.line].axis;
            frame = setup[sg.line].frame;
            tl = pick(sg.c, axis) / g.tile;
            tb0 = sg.lo / g.tile;
            np = sg.hi / g.tile - tb0 + 1;
            item = make_int2((int)i, sg.line);
        }
    }
    // tiles per slot, for the adjoint's partial sums (0 for unused slots)
    if (slot_tiles && i < n_slots) slot_tiles[i] = np;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int p = 0; __ballot(p < np) != 0ull; p++) {
        const bool act = p < np;
        const int key = act ? (int)list_index(tile_index(g, frame, axis, tl, tb0 + p), i) : -1;
        // group the active lanes by tile: one leader (lowest lane) per group
        uint64_t rem = __ballot(act), grp = 0ull;
        int leader = 0;
        while (rem) {
            const int l = __builtin_ctzll(rem);
            const int k = __shfl(key, l, 64);
            const uint64_t m = __ballot(key == k);
            if (key == k) {
                grp = m;
                leader = l;
            }
            rem &= ~m;
        }
        int b = 0;
        if (act && lane == leader) b = atomicAdd(&cur[key], __popcll(grp));
        b = __shfl(b, leader, 64);
        if (act) pairs[tile_off[key] + b + __popcll(grp & below)] = item;
    }
}

// ---------------------------------------------------------------------------
// shared pieces of the tile kernels
// ---------------------------------------------------------------------------
template <typename T>
struct ColRec {
    int32_t s; // short-axis centre, INT_MIN when the column is not emitted
    T Te, Ne;
};

struct ItemHdr {
    int32_t seg;
    int32_t line;
};

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Phase 1: replay a segment's walk from its checkpoint and record, for each
// of its columns, the centre and errors at which the reference emits it, in
// the column slots i * SEG + k of item i of rr.
template <typename T, int SEG>
__device__ __forceinline__ void walk_columns(const LineSetup<T> &L, const Segment<T> &sg, ColRec<T> *rr, int i,
                                             int *flags)
{
    for (int k = 0; k < SEG; k++) rr[i * SEG + k].s = INT_MIN;
    const int a = L.axis, b = 1 - a, sa = pick(L.st, a);
    WalkState<T> s;
    s.c[0] = sg.c[0];
    s.c[1] = sg.c[1];
    s.Te = sg.Te;
    s.Ne = sg.Ne;
    const int c0 = pick(s.c, a);
    int it = 0;
    for (; it < 4 * SEG + 16; it++) {
        if (walk_done(L, s)) break;
        const int k = (pick(s.c, a) - c0) * sa;
        if (k >= sg.ncols) break;
        bool nx, ny;
        walk_next(L, s.Ne, nx, ny);
        if ((a == 0) ? nx : ny) {
            ColRec<T> &r = rr[i * SEG + k];
            r.s = pick(s.c, b);
            r.Te = s.Te;
            r.Ne = s.Ne;
        }
        walk_step(L, s, nx, ny);
    }
    if (it == 4 * SEG + 16) atomicOr(flags, FLAG_STUCK);
}

// Phase 2: the orthogonal fills of one emitted column (bresenham.hpp:364-380),
// restricted to the short-axis window [s_lo, s_hi); emit(s, err) is called
// for every pixel of the window the reference would hand to draw_pixel.
template <typename T, typename Emit>
__device__ __forceinline__ void fill_column(const LineSetup<T> &L, const ColRec<T> &cr, int s_lo, int s_hi,
                                            Emit &&emit)
{
    const int b = 1 - L.axis;
    if (cr.s >= s_lo && cr.s < s_hi) emit(cr.s, cr.Te, cr.Ne);
    const T tb = pick(L.t, b), nb = pick(L.nv, b);
    const int sb = pick(L.st, b);
#pragma unroll
    for (int half = 0; half < 2; half++) {
        const int dir = half ? -sb : sb;
        const T dT = (dir > 0) ? tb : -tb, dN = (dir > 0) ? nb : -nb;
        int s = cr.s;
        T Te = cr.Te, Ne = cr.Ne;
        for (;;) {
            s += dir;
            Te += dT;
            Ne += dN;
            if (!(nsq_of(L, Ne) < L.w2)) break;
            if (s >= s_lo && s < s_hi) emit(s, Te, Ne);
            else if ((dir > 0) ? (s >= s_hi) : (s < s_lo)) break;
        }
    }
}

// LDS accumulator of the tile kernel, by accumulation mode OV:
//   0  sums in double (on gfx950 ds_add_f64 sustains ~20x the lane rate of
//      ds_add_f32; the tile is rounded to the image type once, at the flush)
//   1  max overlap, as order-preserving unsigned codes (ds_max)
//   2  float sums in 32-bit fixed point (ds_add_u32, the fastest LDS atomic,
//      at half the LDS of double; see k_tiles)
template <typename T, int OV>
struct AccType {
    using type = double;
};
template <>
struct AccType<float, 1> {
    using type = uint32_t;
};
template <>
struct AccType<double, 1> {
    using type = unsigned long long;
};
template <>
struct AccType<float, 2> {
    using type = uint32_t;
};

// Adds one pixel value to the LDS tile.  For OV = 2, v is the profile value
// already scaled by the round's 2^F; exact zeros only mark the pixel touched.
// POS: v is known to be positive (and, for OV = 2, at least 1).
template <int OV, bool POS = false, typename Acc, typename V>
__device__ __forceinline__ void tile_put(Acc *img, uint32_t *zb, int idx, V v)
{
    if constexpr (OV == 0) {
        atomicAdd(&img[idx], (Acc)v);
    } else if constexpr (OV == 1) {
        if (v == v) atomicMax(&img[idx], ord_enc(v));
    } else if constexpr (POS) {
        atomicAdd(&img[idx], __float2uint_rn((float)v));
    } else {
        if (v > 0.0f) atomicAdd(&img[idx], max(1u, __float2uint_rn((float)v)));
        else atomicOr(&zb[idx >> 5], 1u << (idx & 31));
    }
}

// ---------------------------------------------------------------------------
// 4. tile kernel (image mode)
// ---------------------------------------------------------------------------
// Per-item record in LDS, written by the phase-1 thread that replays the
// item's walk (it holds the line's setup in registers) and read by the
// phase-2 threads of its columns: the item's line and first long-axis
// coordinate, and for float lines the fast path's constants, so the fills
// need no global load (the exact error near the threshold reads the setup).
struct alignas(16) FillRec {
    float LL, ea, nlim;
    float tb, nb;     // tangent and normal components along the short axis b
    int32_t c0;       // long-axis coordinate of the item's first column
    int32_t line;
    int32_t meta;     // FR_* bits, profile id in bits 8-15
};
constexpr int FR_AXIS = 1, FR_SA = 2, FR_SB = 4, FR_SAFE = 8;

template <typename T>
__device__ __forceinline__ FillRec fill_rec(const LineSetup<T> &L, int c0, int line)
{
    FillRec r;
    const int a = L.axis, b = 1 - a;
    r.c0 = c0;
    r.line = line;
    r.meta = a | (pick(L.st, a) > 0 ? FR_SA : 0) | (pick(L.st, b) > 0 ? FR_SB : 0) |
             (sizeof(T) == 4 && L.safe ? FR_SAFE : 0) | ((int)L.kid << 8);
    r.LL = (float)L.LL;
    r.ea = L.ea;
    r.nlim = L.nlim;
    r.tb = (float)pick(L.t, b);
    r.nb = (float)pick(L.nv, b);
    return r;
}

// The fills of one column half for the float fast path (lines with L.safe):
// the band test |Ne| < nlim, the one-product error estimate (err_fast), the
// profile's constant folded into the column's value scale and the tile window
// turned into a skip count and a trip count, with the LDS index stepped per
// pixel.  The errors advance by the same FP32 additions as the reference's
// loop.  Lg is the line's setup in global memory, read only for exact errors
// near the threshold.  vscale is max_val (OV 0, 1) or the fixed-point scale
// 2^F (OV 2: F >= 20, so positive profiles give values of at least 1).
template <int K, int OV, typename Acc>
__device__ __forceinline__ void fill_half_fast(Acc *img, uint32_t *zb, const FillRec &R,
                                               const LineSetup<float> &Lg, int cs, float Te, float Ne, int half,
                                               int cbase, int sstride, int s_lo, int s_hi, float vscale)
{
    constexpr bool POS = (OV == 2) && fast_positive<K>();
    const float LL = R.LL, ea = R.ea, nlim = R.nlim;
    const int kid = R.meta >> 8;
    const int sb = (R.meta & FR_SB) ? 1 : -1;
    const int dir = half ? -sb : sb;
    const float dT = (dir > 0) ? R.tb : -R.tb, dN = (dir > 0) ? R.nb : -R.nb;
    const float c = fast_scale<K>(kid, vscale);
    auto put = [&](int idx, float T, float N) {
        const float e = err_fast(Lg, LL, ea, T, N);
        if (e <= 1.0f) tile_put<OV, POS>(img, zb, idx, fast_value<K>(kid, e, c));
    };
    wait_lds_loads();
    if (half == 0 && cs >= s_lo && cs < s_hi) put(cbase + (cs - s_lo) * sstride, Te, Ne);
    // steps to the last in-window pixel in the direction of travel, and the
    // first step that lands inside the window
    const int nsteps = (dir > 0) ? (s_hi - 1 - cs) : (cs - s_lo);
    if (nsteps < 1) return;
    const int jin = max(1, (dir > 0) ? (s_lo - cs) : (cs - s_hi + 1));
    for (int j = 1; j < jin; j++) {
        Te += dT;
        Ne += dN;
        if (!(__builtin_fabsf(Ne) < nlim)) return;
    }
    const int step = dir * sstride;
    int idx = cbase + (cs + jin * dir - s_lo) * sstride;
    for (int j = jin; j <= nsteps; j++, idx += step) {
        Te += dT;
        Ne += dN;
        if (!(__builtin_fabsf(Ne) < nlim)) return;
        put(idx, Te, Ne);
    }
}

// Value a fixed-point tile pixel adds to the image: acc * max_val * 2^-F, or
// max_val * 0 (the sign rules of adding exact zeros) if only zeros touched
// it.  has = false: the pixel was not touched.
__device__ __forceinline__ double fix_value(uint32_t acc, const uint32_t *zb, int i, double scale, float max_val,
                                            bool &has)
{
    has = acc != 0u || ((zb[i >> 5] >> (i & 31)) & 1u);
    return acc ? (double)acc * scale : (double)(max_val * 0.0f);
}

typedef float v4f __attribute__((ext_vector_type(4)));

// fix_value for a pixel already in a register, in float: o + (acc 2^-F)
// max_val (or + zero, the value of exact zeros, when only zeros touched it:
// zbit set), or o untouched.  acc 2^-F is exact (p2 = 2^-F); the rounding of
// acc to 24 bits and of the two operations stays below 2e-7 of the pixel's
// value, far inside the fixed-point terms' own error (<= 2^-F max_val each).
__device__ __forceinline__ float fold_fix(float o, uint32_t acc, uint32_t zbit, float p2, float max_val, float zero)
{
    const float add = acc ? ((float)acc * p2) * max_val : zero;
    return (acc | zbit) ? o + add : o;
}

// LDS row pitch of the adjoint's gradient tile, in words: 4 words of padding
// per row.  The lanes of one item read the same short-axis coordinate in 8 or
// 16 consecutive rows (steep lines); with a pitch of TILE (a multiple of the
// bank count) those reads fall in one bank, with TILE + 4 in distinct banks.
// (The tile kernel keeps a pitch of TILE: its padded variants measured slower,
// see k_tiles.)
template <int TILE>
constexpr int tile_pitch() { return TILE + 4; }

// Tile kernel.  The workgroup takes its tile's items ((segment, line) pairs)
// in chunks of CH <= NT: phase 1 gives each of CH threads one item (the walk
// replay, column records into LDS); phase 2 gives each thread one (item,
// column, half) unit of the fills at a time.  The first chunk's loads and
// (256-thread float sum tiles) the tile's current values in `out` are
// requested before the LDS tile is cleared, so their latency overlaps.
//
// Fixed-point sums (OV = 2): an item (one segment in one tile) adds at most
// one value per pixel, so a round of n items adds at most n values of at most
// 2^F with F = clz(n): the sum stays below 2^32.  Each value p > 0 adds
// max(1, rn(p 2^F)) (absolute error <= 2^-F max_val per term, so a touched
// pixel stays non-zero); exact zeros set a bit in zb.  Tiles of more than
// FIX_ROUND items take several rounds, folded into the running pixel values.
constexpr int FIX_ROUND = 4095;

// (at least 5 waves per SIMD: the fastest register budget measured)
template <typename T, int K, int OV, int TILE, int SEG, int NT, int CH, bool FAST>
__global__ void __launch_bounds__(NT, 5)
k_tiles(T *__restrict__ out, Geometry g, const int *__restrict__ tile_off, const int2 *__restrict__ pairs,
        const Segment<T> *__restrict__ segs, const LineSetup<T> *__restrict__ setup, T max_val,
        int *__restrict__ flags)
{
    using Acc = typename AccType<T, OV>::type;
    static_assert(OV != 2 || (FAST && sizeof(T) == 4), "fixed-point sums are the float fast path");
    // LDS pitch TILE: the bank conflicts of the fills (steep lines write one
    // bank from consecutive rows, the two halves of a column share a bank) and
    // of the strided item records cost less than removing them did (a padded
    // pitch and rotated record slots spilled registers at 5 waves per SIMD or
    // dropped to 4)
    constexpr int PITCH = TILE;
    constexpr int LDS_PX = TILE * PITCH;   // LDS words of the tile (pixel (y, x) at y * PITCH + x)
    __shared__ alignas(16) Acc img[LDS_PX];
    __shared__ uint32_t zb[OV == 2 ? LDS_PX / 32 : 1];
    static_assert(CH <= NT, "one phase-1 thread per item");
    __shared__ int32_t rs[CH * SEG + CH];   // + one sink slot per phase-1 lane
    __shared__ T rT[CH * SEG + CH];
    __shared__ T rN[CH * SEG + CH];
    __shared__ FillRec frec[CH];

    const int tile = blockIdx.x;
    // (one round trip for the item range and the error flags: a plan's draw
    // reports errors afterwards and leaves `out` untouched then)
    const int beg = tile_off[tile * LIST_SUBS], end = tile_off[(tile + 1) * LIST_SUBS];
    const int fl = *flags;
    if (fl || beg == end) return;
    const int per_frame = g.ntx * g.nty;
    const int64_t f = tile / per_frame;
    const int r = tile - (int)f * per_frame;
    const int ty = r / g.ntx, tx = r - ty * g.ntx;
    const int X0 = tx * TILE, Y0 = ty * TILE;
    const int X1 = min(X0 + TILE, g.W), Y1 = min(Y0 + TILE, g.H);
    const int tid = threadIdx.x;
    T *fo = out + (f * g.H) * (int64_t)g.W;
    constexpr int V = 4;
    const bool full = (X1 - X0 == TILE) && ((g.W & (V - 1)) == 0) && (((uintptr_t)out & 15) == 0);

    // the tile's current values, for the flush (float sums, 256 threads:
    // 4 float4 per thread), requested first
    constexpr bool PREF = (OV != 1) && sizeof(T) == 4 && NT == 256 && TILE == 64;
    constexpr int PV = PREF ? (TILE * TILE) / (NT * V) : 1;
    float4 pre[PV];
    if constexpr (PREF) {
        if (full) {
#pragma unroll
            for (int j = 0; j < PV; j++) {
                const int i = (tid + j * NT) * V;
                const int y = Y0 + i / TILE, x = X0 + (i % TILE);
                if (y < Y1) {
                    const v4f q = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(fo + (int64_t)y * g.W + x));
                    pre[j] = make_float4(q.x, q.y, q.z, q.w);
                }
            }
        }
    }
    // first chunk's item loads (after the tile's loads, which do not wait on them)
    Segment<T> sg0;
    LineSetup<T> L0;
    if (tid < min(CH, end - beg)) {
        const int2 it = pairs[beg + tid];
        sg0 = segs[it.x];
        L0 = setup[it.y];
    }

    auto clear = [&]() {
        if constexpr (OV != 0 && sizeof(Acc) == 4) {
            // zero words (below every max code / empty fixed-point sum), 16 bytes per store
            static_assert(LDS_PX % 4 == 0, "whole 16-byte stores");
            uint4 *p = reinterpret_cast<uint4 *>(img);
            for (int i = tid; i < LDS_PX / 4; i += NT) p[i] = make_uint4(0u, 0u, 0u, 0u);
        } else {
            for (int i = tid; i < LDS_PX; i += NT) {
                if constexpr (OV == 0) img[i] = Acc(-0.0); // additive identity, keeps -0.0 pixels intact
                else img[i] = Acc(0);                      // below every max code
            }
        }
        if constexpr (OV == 2)
            for (int i = tid; i < LDS_PX / 32; i += NT) zb[i] = 0u;
    };
    clear();

    // rounds [rb, re): one for OV 0 / 1, FIX_ROUND items each for OV 2
    for (int rb = beg, re; rb < end; rb = re) {
        re = (OV == 2) ? rb + min(FIX_ROUND, end - rb) : end;
        const int F = (OV == 2) ? __clz(re - rb) : 0;
        // value scale: max_val, or 2^F for fixed-point sums
        const T vscale = (OV == 2) ? (T)__int_as_float((127 + F) << 23) : max_val;
        if (rb != beg) {
            __syncthreads();
            clear();
        }
        for (int first = rb; first < re; first += CH) {
            const int n = min(CH, re - first);
            // (phase 1 writes only the item records: the first chunk needs no
            // barrier after the clear)
            if (first != beg) __syncthreads();
            // phase 1: replay the walk of one item from its checkpoint and
            // record the centre and errors of every emitted column
            if (tid < n) {
                Segment<T> sg;
                LineSetup<T> L;
                if (first == beg) {
                    sg = sg0;
                    L = L0;
                } else {
                    const int2 it = pairs[first + tid];
                    sg = segs[it.x];
                    L = setup[it.y];
                }
                const int a = L.axis, sa = pick(L.st, a);
                const int c0 = pick(sg.c, a);
                // per-line profiles: the wave's items take their record slots
                // grouped by profile, so the fills of a wave mostly share one
                // (fixed-point sums do not depend on the item order)
                int dst = tid;
                if constexpr (K == K_MIXED) {
                    const int kc = (L.kid >= 0 && L.kid < K_TRIANGULAR) ? L.kid : K_TRIANGULAR;
                    const uint64_t below = (1ull << (tid & 63)) - 1ull;
                    int off = 0, rank = 0;
#pragma unroll
                    for (int k = 0; k <= K_TRIANGULAR; k++) {
                        const uint64_t m = __ballot(kc == k);
                        if (kc == k) rank = off + __popcll(m & below);
                        off += __popcll(m);
                    }
                    dst = (tid & ~63) + rank;
                }
                frec[dst] = fill_rec(L, c0, sg.line);
#pragma unroll
                for (int k = 0; k < SEG; k++) rs[dst * SEG + k] = INT_MIN;
                WalkState<T> s;
                s.c[0] = sg.c[0];
                s.c[1] = sg.c[1];
                s.Te = sg.Te;
                s.Ne = sg.Ne;
                // a loop uniform over the wave: lanes whose walk has ended are
                // masked by predicates and store into their sink slot
                bool act = true;
                for (int it = 0; it < 4 * SEG + 16; it++) {
                    act = act && !walk_done(L, s);
                    const int k = (pick(s.c, a) - c0) * sa;
                    act = act && k < sg.ncols;
                    if (__ballot(act) == 0) break;
                    bool nx, ny;
                    walk_next(L, s.Ne, nx, ny);
                    const bool rec = act && ((a == 0) ? nx : ny);
                    const int slot = rec ? dst * SEG + k : CH * SEG + tid;
                    rs[slot] = pick(s.c, 1 - a);
                    rT[slot] = s.Te;
                    rN[slot] = s.Ne;
                    walk_step(L, s, act && nx, act && ny);
                }
                if (act) atomicOr(flags, FLAG_STUCK);   // still walking after the step budget
            }
            __syncthreads();
            // phase 2: one thread per (item, column, half); the centre goes
            // with the half that steps in +st[b] (bresenham.hpp:364-380)
            for (int u = tid; u < n * SEG * 2; u += NT) {
                const int i = u / (2 * SEG), rem = u - i * (2 * SEG);
                const int k = rem >> 1, half = rem & 1;
                const int slot = i * SEG + k;
                const int cs = rs[slot];
                if (cs == INT_MIN) continue;
                const FillRec R = frec[i];
                const int a = R.meta & FR_AXIS, b = 1 - a;
                const int c = R.c0 + ((R.meta & FR_SA) ? k : -k);
                const int s_lo = (a == 0) ? Y0 : X0, s_hi = (a == 0) ? Y1 : X1;
                // LDS stride of the short axis and offset of the long-axis column
                const int sstride = (a == 0) ? PITCH : 1;
                const int cbase = (a == 0) ? (c - X0) : (c - Y0) * PITCH;
                T Te = rT[slot], Ne = rN[slot];
                if constexpr (FAST && sizeof(T) == 4) {
                    if (R.meta & FR_SAFE) {
                        fill_half_fast<K, OV>(img, zb, R, setup[R.line], cs, Te, Ne, half, cbase, sstride, s_lo,
                                              s_hi, (float)vscale);
                        continue;
                    }
                }
                const LineSetup<T> L = setup[R.line];
                auto put = [&](int s, T Te_, T nsq) {
                    T v;
                    if (!pixel_eval<FAST, K>(L, Te_, nsq, vscale, v)) return;
                    tile_put<OV>(img, zb, cbase + (s - s_lo) * sstride, v);
                };
                if (half == 0 && cs >= s_lo && cs < s_hi) put(cs, Te, nsq_of(L, Ne));
                const int sb = pick(L.st, b);
                const int dir = half ? -sb : sb;
                const T tb = pick(L.t, b), nb = pick(L.nv, b);
                const T dT = (dir > 0) ? tb : -tb, dN = (dir > 0) ? nb : -nb;
                int s = cs;
                for (;;) {
                    s += dir;
                    Te += dT;
                    Ne += dN;
                    const T nsq = nsq_of(L, Ne);
                    if (!(nsq < L.w2)) break;
                    if (s >= s_lo && s < s_hi) put(s, Te, nsq);
                    else if ((dir > 0) ? (s >= s_hi) : (s < s_lo)) break;
                }
            }
        }
        if constexpr (OV == 2) {
            // fold the round into the pixel values (registers or `out`)
            __syncthreads();
            const double scale = ldexp((double)max_val, -F);
            bool has;
            if (PREF && full) {
                // all LDS reads first (4 pixels and their zero bits per read),
                // then branch-free updates
                const float zero = (float)max_val * 0.0f;
                const float p2 = __int_as_float((127 - F) << 23);   // 2^-F, F <= 31
                uint4 acc[PV];
                uint32_t zw[PV];
#pragma unroll
                for (int j = 0; j < PV; j++) {
                    const int i = (tid + j * NT) * V;
                    const int li = (i / TILE) * PITCH + (i % TILE);
                    acc[j] = *reinterpret_cast<const uint4 *>(&img[li]);
                    zw[j] = zb[li >> 5] >> (li & 31);
                }
#pragma unroll
                for (int j = 0; j < PV; j++) {
                    pre[j].x = fold_fix(pre[j].x, acc[j].x, zw[j] & 1u, p2, (float)max_val, zero);
                    pre[j].y = fold_fix(pre[j].y, acc[j].y, zw[j] & 2u, p2, (float)max_val, zero);
                    pre[j].z = fold_fix(pre[j].z, acc[j].z, zw[j] & 4u, p2, (float)max_val, zero);
                    pre[j].w = fold_fix(pre[j].w, acc[j].w, zw[j] & 8u, p2, (float)max_val, zero);
                }
            } else {
                for (int i = tid; i < TILE * TILE; i += NT) {
                    const int y = Y0 + i / TILE, x = X0 + (i % TILE);
                    if (y >= Y1 || x >= X1) continue;
                    const int li = (i / TILE) * PITCH + (i % TILE);
                    const double add = fix_value(img[li], zb, li, scale, (float)max_val, has);
                    T *o = fo + (int64_t)y * g.W + x;
                    if (has) *o = (T)((double)*o + add);
                }
            }
        }
    }
    if constexpr (OV == 2) {
        if (PREF && full) {
#pragma unroll
            for (int j = 0; j < PV; j++) {
                const int i = (tid + j * NT) * V;
                const int y = Y0 + i / TILE, x = X0 + (i % TILE);
                if (y < Y1) {
                    // written once per draw: non-temporal
                    const v4f q = {pre[j].x, pre[j].y, pre[j].z, pre[j].w};
                    __builtin_nontemporal_store(q, reinterpret_cast<v4f *>(fo + (int64_t)y * g.W + x));
                }
            }
        }
        return;
    }
    __syncthreads();

    // flush: one coalesced read-modify-write of the tile, 4 pixels per access
    // where the tile row lies inside the frame
    if constexpr (PREF) {
        if (full) {
#pragma unroll
            for (int j = 0; j < PV; j++) {
                const int i = (tid + j * NT) * V;
                const int y = Y0 + i / TILE, x = X0 + (i % TILE);
                if (y >= Y1) continue;
                const int li = (i / TILE) * PITCH + (i % TILE);
                float4 ov = pre[j];
                ov.x = (float)((double)ov.x + img[li]);
                ov.y = (float)((double)ov.y + img[li + 1]);
                ov.z = (float)((double)ov.z + img[li + 2]);
                ov.w = (float)((double)ov.w + img[li + 3]);
                *reinterpret_cast<float4 *>(fo + (int64_t)y * g.W + x) = ov;
            }
            return;
        }
    }
    for (int i = tid * V; i < TILE * TILE; i += NT * V) {
        const int y = Y0 + i / TILE, x = X0 + (i % TILE);
        if (y >= Y1) continue;
        T *o = fo + (int64_t)y * g.W + x;
        const int li = (i / TILE) * PITCH + (i % TILE);
        if (full) {
            if constexpr (OV == 0) {
                if constexpr (sizeof(T) == 4) {
                    float4 ov = *reinterpret_cast<float4 *>(o);
                    ov.x = (float)((double)ov.x + img[li]);
                    ov.y = (float)((double)ov.y + img[li + 1]);
                    ov.z = (float)((double)ov.z + img[li + 2]);
                    ov.w = (float)((double)ov.w + img[li + 3]);
                    *reinterpret_cast<float4 *>(o) = ov;
                } else {
                    for (int j = 0; j < V; j++) o[j] = o[j] + img[li + j];
                }
            } else {
                for (int j = 0; j < V; j++) {
                    const Acc e = img[li + j];
                    if (e != Acc(0)) {
                        const T m = ord_dec(e);
                        const T ov = o[j];
                        o[j] = (ov < m) ? m : ov; // std::max(out, value)
                    }
                }
            }
        } else {
            for (int j = 0; j < V; j++) {
                if (x + j >= X1) break;
                if constexpr (OV == 0) {
                    o[j] = (T)((double)o[j] + img[li + j]);
                } else {
                    const Acc e = img[li + j];
                    if (e != Acc(0)) {
                        const T m = ord_dec(e);
                        const T ov = o[j];
                        o[j] = (ov < m) ? m : ov;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// emission pass (pixel lists and the deterministic mode)
// ---------------------------------------------------------------------------
template <typename T, int K, bool WRITE>
__global__ void __launch_bounds__(256)
k_emit(const Segment<T> *__restrict__ segs, int64_t n_segs, const LineSetup<T> *__restrict__ setup, int H, int W,
       T max_val, int *__restrict__ cnt, const int64_t *__restrict__ off, uint64_t *__restrict__ keys,
       T *__restrict__ vals, int32_t *__restrict__ lines_out, int *__restrict__ flags)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_segs) return;
    const Segment<T> sg = segs[i];
    const LineSetup<T> L = setup[sg.line];
    const int a = L.axis, b = 1 - a, sa = pick(L.st, a);
    const int db = (b == 0) ? W : H;
    const int64_t base = WRITE ? off[i] : 0;
    int n = 0;

    WalkState<T> s;
    s.c[0] = sg.c[0];
    s.c[1] = sg.c[1];
    s.Te = sg.Te;
    s.Ne = sg.Ne;
    const int c0 = pick(s.c, a);
    int it = 0;
    for (; it < 4 * 64 + 16; it++) {
        if (walk_done(L, s)) break;
        const int k = (pick(s.c, a) - c0) * sa;
        if (k >= sg.ncols) break;
        bool nx, ny;
        walk_next(L, s.Ne, nx, ny);
        if ((a == 0) ? nx : ny) {
            ColRec<T> cr;
            cr.s = pick(s.c, b);
            cr.Te = s.Te;
            cr.Ne = s.Ne;
            const int c = pick(s.c, a);
            fill_column(L, cr, 0, db, [&](int sp, T Te, T Ne) {
                const T err = err_of(L, Te, Ne);
                if (!(err <= T(1))) return;
                if (WRITE) {
                    const int x = (a == 0) ? c : sp, y = (a == 0) ? sp : c;
                    keys[base + n] = ((uint64_t)L.frame * (uint64_t)H + (uint64_t)y) * (uint64_t)W + (uint64_t)x;
                    vals[base + n] = pixel_value<K>(err, max_val);
                    if (lines_out) lines_out[base + n] = sg.line;
                }
                n++;
            });
        }
        walk_step(L, s, nx, ny);
    }
    if (it == 4 * 64 + 16) atomicOr(flags, FLAG_STUCK);
    if (!WRITE) cnt[i] = n;
}

__global__ void k_i32_to_i64(const int *__restrict__ in, int64_t *__restrict__ out, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

template <typename T>
__global__ void k_raw_out(const uint64_t *__restrict__ keys, const T *__restrict__ vals,
                          const int32_t *__restrict__ line_of, int64_t n, int64_t H, int64_t W,
                          int32_t *__restrict__ of, int32_t *__restrict__ oy, int32_t *__restrict__ ox,
                          T *__restrict__ ov, int32_t *__restrict__ ol)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t key = keys[i];
    const uint64_t hw = (uint64_t)H * (uint64_t)W;
    of[i] = (int32_t)(key / hw);
    const uint64_t r = key % hw;
    oy[i] = (int32_t)(r / (uint64_t)W);
    ox[i] = (int32_t)(r % (uint64_t)W);
    ov[i] = vals[i];
    if (ol) ol[i] = line_of[i];
}

// ---------------------------------------------------------------------------
// ordered (reference-order) accumulation: deterministic images and merged
// pixel lists
// ---------------------------------------------------------------------------

// Sorts each tile's item list by segment id, i.e. by (line, walk position):
// the order in which the reference's single thread visits the segments.
// One workgroup per tile, bitonic sort in LDS.
constexpr int SORT_MAX = 8192;
constexpr int FLAG_TILE_FULL = 8;

__global__ void __launch_bounds__(256)
k_sort_tile_items(const int *__restrict__ tile_off, int *__restrict__ items, int *__restrict__ flags,
                  const int *__restrict__ tile_list)
{
    __shared__ int buf[SORT_MAX];
    const int tile = tile_list ? tile_list[blockIdx.x] : (int)blockIdx.x;
    const int beg = tile_off[tile], n = tile_off[tile + 1] - beg;
    if (n <= 1) return;
    if (n > SORT_MAX) {
        if (threadIdx.x == 0) atomicOr(flags, FLAG_TILE_FULL);
        return;
    }
    int p2 = 2;
    while (p2 < n) p2 <<= 1;
    for (int i = threadIdx.x; i < p2; i += blockDim.x) buf[i] = (i < n) ? items[beg + i] : INT_MAX;
    __syncthreads();
    for (int k = 2; k <= p2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < p2; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const int a = buf[i], b = buf[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        buf[i] = b;
                        buf[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) items[beg + i] = buf[i];
}

// One wave per tile.  Items are applied one at a time in sorted order; the
// lanes of an item cover its columns (no pixel repeats inside one line), so
// every pixel sees its contributions in ascending line order, starting from
// the value already in the frame (image mode) or from its first contribution
// (list mode: `first` records the first contributing line).
__device__ inline unsigned int val_bits(float v) { return __float_as_uint(v); }
__device__ inline unsigned long long val_bits(double v) { return (unsigned long long)__double_as_longlong(v); }
__device__ inline float bits_val(unsigned int b, float) { return __uint_as_float(b); }
__device__ inline double bits_val(unsigned long long b, double) { return __longlong_as_double((long long)b); }

// *p = max(*p, v) in the order std::max(*p, v) takes: replaced only if *p < v
__device__ inline void atomic_max_val(float *p, float v)
{
    unsigned int *q = (unsigned int *)p;
    unsigned int old = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (bits_val(old, v) < v) {
        const unsigned int prev = atomicCAS(q, old, val_bits(v));
        if (prev == old) break;
        old = prev;
    }
}

__device__ inline void atomic_max_val(double *p, double v)
{
    unsigned long long *q = (unsigned long long *)p;
    unsigned long long old = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (bits_val(old, v) < v) {
        const unsigned long long prev = atomicCAS(q, old, val_bits(v));
        if (prev == old) break;
        old = prev;
    }
}

// ACC >= 0 (accumulate_lines): the plan's frames are terms; every touched
// pixel of a term tile (accumulated from zero, as with LIST) is combined into
// frame term_frame[term] of img with atomics, sum (ACC 0) or max (ACC 1).
template <typename T, int K, int OV, int TILE, int SEG, bool LIST, int ACC>
__global__ void __launch_bounds__(64)
k_tiles_ordered(T *__restrict__ img, uint8_t *__restrict__ touched, int32_t *__restrict__ first_line, Geometry g,
                const int *__restrict__ tile_off, const int *__restrict__ items, const Segment<T> *__restrict__ segs,
                const LineSetup<T> *__restrict__ setup, T max_val, int *__restrict__ flags,
                const int32_t *__restrict__ term_frame, const int *__restrict__ tile_list)
{
    static_assert(ACC < 0 || LIST, "accumulation starts from zero");
    static_assert(2 * SEG <= 64, "two lanes per column");
    __shared__ T acc[TILE * TILE];
    __shared__ int32_t first[TILE * TILE];
    __shared__ ColRec<T> rec[64 * SEG];
    __shared__ ItemHdr hdr[64];

    const int tile = tile_list ? tile_list[blockIdx.x] : (int)blockIdx.x;
    const int beg = tile_off[tile], end = tile_off[tile + 1];
    // a tile too large to sort (k_sort_tile_items, earlier on the stream)
    // fails the call: leave every output untouched
    if (beg == end || (__hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & FLAG_TILE_FULL))
        return;
    const int per_frame = g.ntx * g.nty;
    const int64_t f = tile / per_frame;
    const int r = tile - (int)f * per_frame;
    const int ty = r / g.ntx, tx = r - ty * g.ntx;
    const int X0 = tx * TILE, Y0 = ty * TILE;
    const int X1 = min(X0 + TILE, g.W), Y1 = min(Y0 + TILE, g.H);
    T *fo = ACC >= 0 ? img : img + (f * g.H) * (int64_t)g.W;
    const int lane = threadIdx.x;

    for (int i = lane; i < TILE * TILE; i += 64) {
        const int y = Y0 + i / TILE, x = X0 + (i % TILE);
        if (LIST) acc[i] = T(0);
        else acc[i] = (y < Y1 && x < X1) ? fo[(int64_t)y * g.W + x] : T(0);
        first[i] = -1;
    }
    wave_sync();

    for (int c0 = beg; c0 < end; c0 += 64) {
        const int n = min(64, end - c0);
        if (lane < n) {
            const int sid = items[c0 + lane];
            const Segment<T> sg = segs[sid];
            const LineSetup<T> L = setup[sg.line];
            hdr[lane].seg = sid;
            hdr[lane].line = sg.line;
            walk_columns<T, SEG>(L, sg, rec, lane, flags);
        }
        wave_sync();
        for (int j = 0; j < n; j++) {
            if (lane < 2 * SEG) {
                const int k = lane >> 1, half = lane & 1;
                const ColRec<T> cr = rec[j * SEG + k];
                if (cr.s != INT_MIN) {
                    const Segment<T> &sg = segs[hdr[j].seg];
                    const LineSetup<T> &L = setup[hdr[j].line];
                    const int line = hdr[j].line;
                    const int a = L.axis, b = 1 - a;
                    const int c = pick(sg.c, a) + k * pick(L.st, a);
                    const int s_lo = (a == 0) ? Y0 : X0, s_hi = (a == 0) ? Y1 : X1;
                    auto put = [&](int s, T Te, T Ne) {
                        const T err = err_of(L, Te, Ne);
                        if (!(err <= T(1))) return;
                        const int x = (a == 0) ? c : s, y = (a == 0) ? s : c;
                        const int p = (y - Y0) * TILE + (x - X0);
                        const T v = pixel_value<K>(err, max_val);
                        if (LIST && first[p] < 0) {
                            first[p] = line;
                            // accumulate_lines' term buffer starts at zero: 0 + v, or
                            // std::max(0, v) (bresenham.cpp:218-221)
                            acc[p] = (ACC < 0) ? v : ((OV == 0) ? T(0) + v : ((T(0) < v) ? v : T(0)));
                        } else {
                            if (OV == 0) acc[p] = acc[p] + v;
                            else acc[p] = (acc[p] < v) ? v : acc[p];
                            if (!LIST && first[p] < 0) first[p] = line;
                        }
                    };
                    // lane pair (k, 0): centre and the +step half; (k, 1): the -step half
                    if (half == 0 && cr.s >= s_lo && cr.s < s_hi) put(cr.s, cr.Te, cr.Ne);
                    const int dir = half ? -pick(L.st, b) : pick(L.st, b);
                    const T dT = (dir > 0) ? pick(L.t, b) : -pick(L.t, b), dN = (dir > 0) ? pick(L.nv, b) : -pick(L.nv, b);
                    int s = cr.s;
                    T Te = cr.Te, Ne = cr.Ne;
                    for (;;) {
                        s += dir;
                        Te += dT;
                        Ne += dN;
                        if (!(nsq_of(L, Ne) < L.w2)) break;
                        if (s >= s_lo && s < s_hi) put(s, Te, Ne);
                        else if ((dir > 0) ? (s >= s_hi) : (s < s_lo)) break;
                    }
                }
            }
            wave_sync();
        }
    }

    for (int i = lane; i < TILE * TILE; i += 64) {
        const int y = Y0 + i / TILE, x = X0 + (i % TILE);
        if (y >= Y1 || x >= X1) continue;
        const int64_t o = (int64_t)y * g.W + x;
        if (ACC >= 0) {
            if (first[i] >= 0) {
                T *dst = img + ((int64_t)term_frame[f] * g.H) * (int64_t)g.W + o;
                if (ACC == 0) atomicAdd(dst, acc[i]);
                else atomic_max_val(dst, acc[i]);
            }
        } else if (LIST) {
            const int64_t q = (f * g.H) * (int64_t)g.W + o;
            touched[q] = first[i] >= 0;
            if (first[i] >= 0) {
                fo[o] = acc[i];
                if (first_line) first_line[q] = first[i];
            }
        } else if (first[i] >= 0) {
            fo[o] = acc[i];
        }
    }
}

// stream compaction of a dense (image, touched) pair in raster order
__global__ void k_row_count(const uint8_t *__restrict__ touched, int64_t rows, int W, int64_t *__restrict__ cnt)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const uint8_t *t = touched + r * W;
    int64_t c = 0;
    for (int x = 0; x < W; x++) c += t[x];
    cnt[r] = c;
}

template <typename T>
__global__ void k_row_write(const uint8_t *__restrict__ touched, const T *__restrict__ img,
                            const int32_t *__restrict__ first_line, int64_t rows, int H, int W,
                            const int64_t *__restrict__ off, int32_t *__restrict__ of, int32_t *__restrict__ oy,
                            int32_t *__restrict__ ox, T *__restrict__ ov, int32_t *__restrict__ ol)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    int64_t p = off[r];
    const int64_t base = r * W;
    for (int x = 0; x < W; x++) {
        if (!touched[base + x]) continue;
        of[p] = (int32_t)(r / H);
        oy[p] = (int32_t)(r % H);
        ox[p] = x;
        ov[p] = img[base + x];
        if (ol) ol[p] = first_line[base + x];
        p++;
    }
}

// ---------------------------------------------------------------------------
// adjoint (sum overlap): per (segment, tile) partial gradients, then an
// ordered per-line reduction
// ---------------------------------------------------------------------------
// Per-line constants of the per-pixel gradient terms (double).  The tangent
// is the exact difference of the end points (the walk's float tangent
// p1 - p0 is rounded, which moves dist by ~1e-5 relative at these
// magnitudes); cc and cw fold max_val, the profile's constant and the powers
// of 1/w of the two gradient factors (see pixel_grad).
struct GradConst {
    double x0, y0, tx, ty, inv_mag, inv_w2, cc, cw;
};

template <int K, typename T>
__device__ __forceinline__ GradConst grad_const(const LineSetup<T> &L, const T *line, T max_val)
{
    GradConst c;
    c.x0 = (double)L.p0[0];
    c.y0 = (double)L.p0[1];
    c.tx = (double)line[2] - c.x0;
    c.ty = (double)line[3] - c.y0;
    c.inv_mag = 1.0 / (c.tx * c.tx + c.ty * c.ty);
    const double inv_w = 1.0 / (double)L.w;
    c.inv_w2 = inv_w * inv_w;
    const double mv = (double)max_val;
    // K(sqrt(u)) = k(u): dK/d(dist^2) = k'(u) / w^2 and dK/dw = k'(u) (-2 u / w);
    // k'(u) = -1.875 (1 - u) (biweight), -0.75 (parabolic), -4.5 k(u) (gaussian)
    double kp;
    if constexpr (K == K_BIWEIGHT) kp = -1.875;
    else if constexpr (K == K_PARABOLIC) kp = -0.75;
    else if constexpr (K == K_GAUSSIAN) kp = -4.5 * 0.3989422804014327;
    else kp = 0.0;
    if constexpr (K == K_TRIANGULAR) {
        // K = 1 - dist / w: dK/d(dist^2) = -1 / (2 w dist), dK/dw = dist / w^2
        c.cc = mv * inv_w;
        c.cw = mv * c.inv_w2;
    } else {
        c.cc = -2.0 * kp * mv * c.inv_w2;
        c.cw = -2.0 * kp * mv * inv_w;
    }
    return c;
}

// Per-pixel gradient terms of max_val K(dist / w), in double.  With q the
// clamped projection of p (parameter ts in [0, 1]) and (dx, dy) = p - q,
// d(dist^2)/d(x0, y0) = -2 (1 - ts) (dx, dy) and d(dist^2)/d(x1, y1) =
// -2 ts (dx, dy).  The accumulators hold
//   acc[0..1] = sum c (dx, dy),  acc[2..3] = sum c ts (dx, dy),  acc[4] = dL/dw,
// with c = -2 g max_val dK/d(dist^2); the end-point gradients are then
// (acc[0] - acc[2], acc[1] - acc[3], acc[2], acc[3]) (grad_finish).
template <int K>
__device__ __forceinline__ void pixel_grad(const GradConst &C, int x, int y, double gv, double *acc)
{
    const double rx = (double)x - C.x0, ry = (double)y - C.y0;
    double ts = __builtin_fma(ry, C.ty, rx * C.tx) * C.inv_mag;
    ts = (ts < 0.0) ? 0.0 : ((ts > 1.0) ? 1.0 : ts);
    const double dx = __builtin_fma(-ts, C.tx, rx), dy = __builtin_fma(-ts, C.ty, ry);
    const double d2 = __builtin_fma(dx, dx, dy * dy);
    double c, gw;
    if constexpr (K == K_TRIANGULAR) {
        const double dist = sqrt(d2);
        c = (dist > 0.0) ? gv * C.cc / dist : 0.0;
        gw = gv * dist;
    } else {
        const double u = d2 * C.inv_w2;
        double f;   // k'(u) / k'(0)
        if constexpr (K == K_BIWEIGHT) f = 1.0 - u;
        else if constexpr (K == K_GAUSSIAN) f = ::exp(-4.5 * u);
        else f = 1.0;
        const double gf = gv * f;
        c = gf * C.cc;
        gw = gf * u;
    }
    const double cts = c * ts;
    acc[0] = __builtin_fma(c, dx, acc[0]);
    acc[1] = __builtin_fma(c, dy, acc[1]);
    acc[2] = __builtin_fma(cts, dx, acc[2]);
    acc[3] = __builtin_fma(cts, dy, acc[3]);
    acc[4] = __builtin_fma(gw, C.cw, acc[4]);
}

// The fills of one column for the float fast path (both halves and the
// centre), with the forward's band test and inclusion decisions (err_fast);
// lines with L.safe only.
template <int K, int TILE>
__device__ __forceinline__ void adj_column_fast(const LineSetup<float> &L, const GradConst &C, const float *gt, int cs,
                                                float Te0, float Ne0, int c, int X0, int Y0, int s_lo, int s_hi,
                                                double *acc)
{
    const float LL = L.LL, ea = L.ea, nlim = L.nlim;
    const int a = L.axis, b = 1 - a;
    auto visit = [&](int s, float T, float N) {
        if (!(err_fast(L, LL, ea, T, N) <= 1.0f)) return;
        const int x = (a == 0) ? c : s, y = (a == 0) ? s : c;
        const float gv = gt[(y - Y0) * tile_pitch<TILE>() + (x - X0)];
        if (gv != 0.0f) pixel_grad<K>(C, x, y, (double)gv, acc);
    };
    if (cs >= s_lo && cs < s_hi) visit(cs, Te0, Ne0);
    const int sb = pick(L.st, b);
#pragma unroll
    for (int half = 0; half < 2; half++) {
        const int dir = half ? -sb : sb;
        const float dT = (dir > 0) ? pick(L.t, b) : -pick(L.t, b), dN = (dir > 0) ? pick(L.nv, b) : -pick(L.nv, b);
        const int nsteps = (dir > 0) ? (s_hi - 1 - cs) : (cs - s_lo);
        const int jin = (dir > 0) ? (s_lo - cs) : (cs - s_hi + 1);
        float Te = Te0, Ne = Ne0;
        for (int j = 1; j <= nsteps; j++) {
            Te += dT;
            Ne += dN;
            if (!(__builtin_fabsf(Ne) < nlim)) break;
            if (j >= jin) visit(cs + j * dir, Te, Ne);
        }
    }
}

template <typename T, int K, int TILE, int SEG, int NT, int CH>
__global__ void __launch_bounds__(NT)
k_adjoint_tiles(const T *__restrict__ grad, const T *__restrict__ lines, Geometry g, const int *__restrict__ tile_off,
                const int2 *__restrict__ pairs, const Segment<T> *__restrict__ segs,
                const LineSetup<T> *__restrict__ setup, const int *__restrict__ slot_base, T max_val,
                double *__restrict__ partial, int *__restrict__ flags)
{
    constexpr int NW = NT / 64;
    static_assert(64 % SEG == 0, "a segment's columns must share one round of a wave");
    __shared__ T gt[TILE * tile_pitch<TILE>()];
    __shared__ ColRec<T> rec[NW][CH * SEG];
    __shared__ ItemHdr hdr[NW][CH];

    const int tile = blockIdx.x;
    const int beg = tile_off[tile * LIST_SUBS], end = tile_off[(tile + 1) * LIST_SUBS];
    // a plan whose last draw failed (capacity, bad idxs) has no valid lists
    if (*flags || beg == end) return;
    const int per_frame = g.ntx * g.nty;
    const int64_t f = tile / per_frame;
    const int r = tile - (int)f * per_frame;
    const int ty = r / g.ntx, tx = r - ty * g.ntx;
    const int X0 = tx * TILE, Y0 = ty * TILE;
    const int X1 = min(X0 + TILE, g.W), Y1 = min(Y0 + TILE, g.H);

    const T *fg = grad + (f * g.H) * (int64_t)g.W;
    // the tile of dLoss/dImage: every load issued before the LDS writes
    static_assert((TILE * TILE) % NT == 0, "whole tile rows per thread pass");
    constexpr int PER = TILE * TILE / NT;
    T gv[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int i = threadIdx.x + j * NT;
        const int y = Y0 + i / TILE, x = X0 + (i % TILE);
        gv[j] = (y < Y1 && x < X1) ? __builtin_nontemporal_load(&fg[(int64_t)y * g.W + x]) : T(0);
    }
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int i = threadIdx.x + j * NT;
        gt[(i / TILE) * tile_pitch<TILE>() + (i % TILE)] = gv[j];
    }
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    ColRec<T> *R = rec[wave];
    for (int first = beg + wave * CH; first < end; first += NW * CH) {
        const int n = min(CH, end - first);
        if (lane < n) {
            const int2 it = pairs[first + lane];
            const Segment<T> sg = segs[it.x];
            const LineSetup<T> L = setup[it.y];
            hdr[wave][lane].seg = it.x;
            hdr[wave][lane].line = it.y;
            walk_columns<T, SEG>(L, sg, R, lane, flags);
        }
        wave_sync();
        const int rounds = (n * SEG + 63) / 64;
        for (int rd = 0; rd < rounds; rd++) {
            const int u = lane + 64 * rd;
            const int i = u / SEG, k = u - i * SEG;
            double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            if (i < n) {
                const ColRec<T> cr = R[u];
                if (cr.s != INT_MIN) {
                    const Segment<T> &sg = segs[hdr[wave][i].seg];
                    const LineSetup<T> &L = setup[hdr[wave][i].line];
                    const int a = L.axis;
                    const int c = pick(sg.c, a) + k * pick(L.st, a);
                    const int s_lo = (a == 0) ? Y0 : X0, s_hi = (a == 0) ? Y1 : X1;
                    const GradConst C = grad_const<K>(L, lines + 5 * (int64_t)hdr[wave][i].line, max_val);
                    bool done = false;
                    if constexpr (sizeof(T) == 4) {
                        if (L.safe) {
                            adj_column_fast<K, TILE>(L, C, (const float *)gt, cr.s, cr.Te, cr.Ne, c, X0, Y0, s_lo,
                                                     s_hi, acc);
                            done = true;
                        }
                    }
                    if (!done) {
                        fill_column(L, cr, s_lo, s_hi, [&](int s, T Te, T Ne) {
                            const T err = err_of(L, Te, Ne);
                            if (!(err <= T(1))) return;
                            const int x = (a == 0) ? c : s, y = (a == 0) ? s : c;
                            const T gv = gt[(y - Y0) * tile_pitch<TILE>() + (x - X0)];
                            if (gv != T(0)) pixel_grad<K>(C, x, y, (double)gv, acc);
                        });
                    }
                }
            }
            // the SEG lanes of one item hold its columns: reduce them in a fixed tree
#pragma unroll
            for (int m = SEG / 2; m > 0; m >>= 1)
#pragma unroll
                for (int j = 0; j < 5; j++) acc[j] += __shfl_xor(acc[j], m, SEG);
            if (k == 0 && i < n) {
                const int sid = hdr[wave][i].seg;
                const Segment<T> &sg = segs[sid];
                const int a = setup[hdr[wave][i].line].axis;
                const int tb = (a == 0) ? ty : tx;
                double *p = partial + 5 * (int64_t)(slot_base[sid] + (tb - sg.lo / g.tile));
#pragma unroll
                for (int j = 0; j < 5; j++) p[j] = acc[j];
            }
        }
        wave_sync();
    }
}

// Per-line sums of the partials (deterministic): LG_LANES lanes per line,
// lane j adding the line's partials j, j + LG_LANES, ... in slot order, then a
// fixed butterfly over the lanes.  A line's slots are [line_base[l],
// line_base[l] + segment_bound); slot s's partials start at slot_base[s].
constexpr int LG_LANES = 8;

template <typename T>
__global__ void __launch_bounds__(256)
k_line_grad(const int64_t *__restrict__ line_base, const LineSetup<T> *__restrict__ setup, Geometry g,
            const int *__restrict__ slot_base, const double *__restrict__ partial, int64_t n_lines,
            T *__restrict__ grad_lines, const int *__restrict__ flags)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t l = t / LG_LANES;
    const int j = (int)(t % LG_LANES);
    // after a failed draw of a plan the slot ranges are not valid (the
    // gradient stays zero; the error is reported by the plan's check)
    const bool ok = l < n_lines && *flags == 0;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (ok) {
        const int ub = segment_bound(setup[l], g);
        const int64_t s0 = slot_base[line_base[l]], s1 = slot_base[line_base[l] + ub];
        for (int64_t s = s0 + j; s < s1; s += LG_LANES)
#pragma unroll
            for (int k = 0; k < 5; k++) acc[k] += partial[5 * s + k];
    }
#pragma unroll
    for (int m = LG_LANES / 2; m > 0; m >>= 1)
#pragma unroll
        for (int k = 0; k < 5; k++) acc[k] += __shfl_xor(acc[k], m, LG_LANES);
    if (!ok || j != 0) return;
    // (sum c dx - sum c ts dx, ..., sum c ts dx, ...): see pixel_grad
    grad_lines[5 * l + 0] = (T)(acc[0] - acc[2]);
    grad_lines[5 * l + 1] = (T)(acc[1] - acc[3]);
    grad_lines[5 * l + 2] = (T)acc[2];
    grad_lines[5 * l + 3] = (T)acc[3];
    grad_lines[5 * l + 4] = (T)acc[4];
}

// ---------------------------------------------------------------------------
// host pipeline
// ---------------------------------------------------------------------------
template <typename T>
struct SegmentPlan {
    Geometry g;
    int64_t n_lines = 0;
    int64_t n_segs = 0;
    int64_t n_items = 0;
    int64_t n_tiles = 0;
    DevBuf<LineSetup<T>> setup;
    DevBuf<int> seg_cnt, seg_off, tile_cnt, tile_off, tile_cur, items, flags;
    DevBuf<int2> pairs; // unordered plan: per-tile (segment, line) lists
    DevBuf<int> slot_tiles; // unordered plan, when asked for: tiles per segment slot (the adjoint's)
    DevBuf<Segment<T>> segs;
};

static int check_common(int64_t n_frames, int64_t H, int64_t W, int64_t n_lines, bool have_idxs, int kernel,
                        int overlap)
{
    if (kernel < 0 || kernel > 4) return fail(CBX_EVALUE, "kernel is missing for id " + std::to_string(kernel));
    if (overlap < 0 || overlap > 1) return fail(CBX_EVALUE, "Invalid overlap keyword");
    if (n_frames < 0 || H < 0 || W < 0 || n_lines < 0) return fail(CBX_EVALUE, "negative size");
    if (H > (1 << 30) || W > (1 << 30) || n_frames >= (int64_t(1) << 31))
        return fail(CBX_EVALUE, "frame shape exceeds the supported range");
    if (n_lines >= (int64_t(1) << 31)) return fail(CBX_EVALUE, "too many lines");
    if (!have_idxs && n_lines && n_frames != 1 && n_frames != n_lines)
        return fail(CBX_EVALUE, "idxs has an incompatible size (" + std::to_string(n_lines) +
                                    " != " + std::to_string(n_frames) + ")");
    return CBX_OK;
}

static int flags_to_status(int flags)
{
    if (flags & FLAG_INDEX) return fail(CBX_EINDEX, "idxs is outside of (0, n_frames)");
    if (flags & FLAG_LONG)
        return fail(CBX_EVALUE, "a line's walk exceeds the supported length (2^26 px) or coordinate range (1e9)");
    if (flags & FLAG_KERNEL) return fail(CBX_EVALUE, "kernel is missing for a per-line profile id");
    if (flags & FLAG_STUCK) return fail(CBX_ERUNTIME, "a line walk did not terminate within its step budget");
    if (flags & FLAG_OVERFLOW) return fail(CBX_ECAPACITY, "the plan's segment or tile-item capacity was exceeded");
    return CBX_OK;
}

// Builds segments (and, with bin, the per-tile item lists).  One host sync.
template <typename T, typename I>
static int build_plan(SegmentPlan<T> &P, const T *lines, const I *idxs, int kernel, int64_t n_lines,
                      int64_t n_frames, int64_t H, int64_t W, int tile, int seg, bool bin, hipStream_t st,
                      bool skip_neg = false)
{
    Geometry &g = P.g;
    g.skip_neg = skip_neg ? 1 : 0;
    g.n_frames = n_frames;
    g.H = (int32_t)H;
    g.W = (int32_t)W;
    g.tile = tile;
    g.seg = seg;
    g.ntx = (int32_t)((W + tile - 1) / tile);
    g.nty = (int32_t)((H + tile - 1) / tile);
    P.n_lines = n_lines;
    P.n_tiles = bin ? n_frames * g.ntx * g.nty : 0;
    if (P.n_tiles >= (int64_t(1) << 31)) return fail(CBX_EVALUE, "too many tiles");

    CBX_HIP(P.setup.alloc(n_lines, st));
    CBX_HIP(P.seg_cnt.alloc(n_lines + 1, st));
    CBX_HIP(P.seg_off.alloc(n_lines + 1, st));
    CBX_HIP(P.flags.alloc(1, st));
    CBX_HIP(hipMemsetAsync(P.flags.ptr, 0, sizeof(int), st));
    CBX_HIP(hipMemsetAsync(P.seg_cnt.ptr + n_lines, 0, sizeof(int), st));
    if (bin) {
        CBX_HIP(P.tile_cnt.alloc(P.n_tiles + 1, st));
        CBX_HIP(P.tile_off.alloc(P.n_tiles + 1, st));
        CBX_HIP(hipMemsetAsync(P.tile_cnt.ptr, 0, sizeof(int) * (P.n_tiles + 1), st));
    }
    const int blk = 256;
    {
        KernelTimer timer1("segments_pass1", st);
    if (bin)
        k_segments<T, I, false, true><<<grid_for(n_lines, blk), blk, 0, st>>>(
            lines, idxs, kernel, n_lines, g, P.setup.ptr, P.seg_cnt.ptr, P.tile_cnt.ptr, nullptr, nullptr, nullptr, nullptr,
            nullptr, P.flags.ptr);
    else
        k_segments<T, I, false, false><<<grid_for(n_lines, blk), blk, 0, st>>>(
            lines, idxs, kernel, n_lines, g, P.setup.ptr, P.seg_cnt.ptr, nullptr, nullptr, nullptr, nullptr, nullptr,
            nullptr, P.flags.ptr);
    }
    CBX_HIP(hipGetLastError());
    CBX_HIP(exclusive_scan_i32(P.seg_cnt.ptr, P.seg_off.ptr, n_lines + 1, st));
    if (bin) CBX_HIP(exclusive_scan_i32(P.tile_cnt.ptr, P.tile_off.ptr, P.n_tiles + 1, st));

    int host[3] = {0, 0, 0};
    {
        const void *src[3] = {P.flags.ptr, P.seg_off.ptr + n_lines, bin ? (const void *)(P.tile_off.ptr + P.n_tiles)
                                                                         : (const void *)P.flags.ptr};
        const int nb[3] = {4, 4, 4};
        uint64_t hv[3];
        CBX_HIP(read_device_values(st, 3, src, nb, hv));
        for (int i = 0; i < 3; i++) host[i] = (int)(uint32_t)hv[i];
        if (!bin) host[2] = 0;
    }
    int rc = flags_to_status(host[0]);
    if (rc) return rc;
    if (host[1] < 0 || host[2] < 0) return fail(CBX_ERUNTIME, "segment count overflow");
    P.n_segs = host[1];
    P.n_items = host[2];
    if (P.n_segs == 0) return CBX_OK;

    CBX_HIP(P.segs.alloc(P.n_segs, st));
    if (bin) {
        CBX_HIP(P.items.alloc(P.n_items, st));
        CBX_HIP(P.tile_cur.alloc(P.n_tiles, st));
        CBX_HIP(hipMemsetAsync(P.tile_cur.ptr, 0, sizeof(int) * P.n_tiles, st));
        k_segments<T, I, true, true><<<grid_for(n_lines, blk), blk, 0, st>>>(
            lines, idxs, kernel, n_lines, g, P.setup.ptr, nullptr, nullptr, P.seg_off.ptr, P.segs.ptr, P.tile_off.ptr,
            P.tile_cur.ptr, P.items.ptr, P.flags.ptr);
    } else {
        k_segments<T, I, true, false><<<grid_for(n_lines, blk), blk, 0, st>>>(
            lines, idxs, kernel, n_lines, g, P.setup.ptr, nullptr, nullptr, P.seg_off.ptr, P.segs.ptr, nullptr, nullptr,
            nullptr, P.flags.ptr);
    }
    CBX_HIP(hipGetLastError());
    return CBX_OK;
}

// Largest slot total of the eager unordered plans so far (a capacity hint
// for the next).
static std::atomic<int64_t> g_slot_hint{0};

// Unordered plan for the tile path: one walk per line (k_setup_lines,
// k_walk_lines, k_bin_slots).  Segments live in slots (n_segs = slots; unused
// slots have ncols = 0 and are in no item list).  One host read after the
// walk (the item total, flags, the slot total), one more before it when the
// previous call's slot total is no usable capacity.
template <typename T, typename I>
static int build_plan_unordered(SegmentPlan<T> &P, const T *lines, const I *idxs, const uint8_t *kernels,
                                int kernel, int64_t n_lines, int64_t n_frames, int64_t H, int64_t W, int tile,
                                int seg, hipStream_t st, int64_t *line_base = nullptr, bool slot_tiles = false)
{
    Geometry &g = P.g;
    g.n_frames = n_frames;
    g.H = (int32_t)H;
    g.W = (int32_t)W;
    g.tile = tile;
    g.seg = seg;
    g.ntx = (int32_t)((W + tile - 1) / tile);
    g.nty = (int32_t)((H + tile - 1) / tile);
    P.n_lines = n_lines;
    P.n_tiles = n_frames * g.ntx * g.nty;
    if (P.n_tiles >= (int64_t(1) << 31) / LIST_SUBS) return fail(CBX_EVALUE, "too many tiles");
    const int64_t n_lists = P.n_tiles * LIST_SUBS;

    DevBuf<unsigned long long> counters;
    CBX_HIP(P.setup.alloc(n_lines, st));
    CBX_HIP(P.flags.alloc(1, st));
    CBX_HIP(counters.alloc(2, st));
    CBX_HIP(P.tile_cnt.alloc(n_lists, st));
    CBX_HIP(P.tile_off.alloc(n_lists + 1, st));
    CBX_HIP(P.tile_cur.alloc(n_lists, st));
    {
        // every counter of the pass, cleared by one kernel
        void *const zp[4] = {P.flags.ptr, counters.ptr, P.tile_cnt.ptr, P.tile_cur.ptr};
        const int64_t zw[4] = {1, 4, n_lists, n_lists};
        CBX_HIP(zero_words(st, 4, zp, zw));
    }
    const int blk = 256;
    k_setup_lines<T, I><<<grid_for(n_lines, blk), blk, 0, st>>>(lines, idxs, kernels, kernel, n_lines, g,
                                                               P.setup.ptr, counters.ptr, P.flags.ptr);
    CBX_HIP(hipGetLastError());
    uint64_t hv[3];
    int fl = 0, rc = 0;
    // Slot capacity: the largest slot total of the previous calls when it is
    // not far above what these lines could need, so the walk starts without
    // reading this call's total first; a walk that runs out of slots
    // (FLAG_OVERFLOW with a larger total) is cleared and run again at the
    // exact total.  Without a usable hint the total is read first.
    int64_t cap = g_slot_hint.load(std::memory_order_relaxed);
    if (cap <= 0 || cap > 64 * n_lines + 65536) {
        const void *src[2] = {counters.ptr, P.flags.ptr};
        const int nb[2] = {8, 4};
        CBX_HIP(read_device_values(st, 2, src, nb, hv));
        fl = (int)(uint32_t)hv[1];
        if ((rc = flags_to_status(fl))) return rc;
        if (hv[0] >= (1ull << 31)) return fail(CBX_EVALUE, "too many segments (2^31)");
        cap = (int64_t)hv[0];
        if (cap == 0) {
            P.n_segs = 0;
            return CBX_OK;
        }
    }
    int64_t slots = 0;
    for (int attempt = 0;; attempt++) {
        CBX_HIP(P.segs.alloc(cap, st));
        {
            KernelTimer timer("segments_walk", st);
            k_walk_lines<T><<<grid_for(n_lines, blk), blk, 0, st>>>(n_lines, g, P.setup.ptr, P.segs.ptr, cap,
                                                                    counters.ptr, P.tile_cnt.ptr, P.flags.ptr,
                                                                    line_base);
        }
        CBX_HIP(hipGetLastError());
        // list offsets (tile_off[n_lists]: the item total)
        CBX_HIP(exclusive_scan_total_i32(P.tile_cnt.ptr, nullptr, P.tile_off.ptr, n_lists, st));
        {
            const void *src[3] = {P.tile_off.ptr + n_lists, P.flags.ptr, counters.ptr};
            const int nb[3] = {4, 4, 8};
            CBX_HIP(read_device_values(st, 3, src, nb, hv));
        }
        fl = (int)(uint32_t)hv[1];
        if ((rc = flags_to_status(fl & ~FLAG_OVERFLOW))) return rc;
        if (hv[2] >= (1ull << 31)) return fail(CBX_EVALUE, "too many segments (2^31)");
        slots = (int64_t)hv[2];
        if (!(fl & FLAG_OVERFLOW)) break;
        if (attempt > 0 || slots <= cap) return flags_to_status(fl);   // cannot happen: the total fits
        // too few slots: clear what the walk counted and walk again
        cap = slots;
        void *const zp[4] = {P.flags.ptr, counters.ptr + 1, P.tile_cnt.ptr, P.tile_cur.ptr};
        const int64_t zw[4] = {1, 2, n_lists, n_lists};
        CBX_HIP(zero_words(st, 4, zp, zw));
    }
    // the largest total so far: calls of varying sizes rarely walk twice
    for (int64_t h = g_slot_hint.load(std::memory_order_relaxed);
         h < slots && !g_slot_hint.compare_exchange_weak(h, slots, std::memory_order_relaxed);) {
    }
    P.n_segs = slots;
    if (P.n_segs == 0) return CBX_OK;
    const int n_items = (int)(uint32_t)hv[0];
    if (n_items < 0) return fail(CBX_ERUNTIME, "tile item count overflow");
    P.n_items = n_items;
    CBX_HIP(P.pairs.alloc(P.n_items, st));
    if (slot_tiles) CBX_HIP(P.slot_tiles.alloc(P.n_segs, st));
    k_bin_slots<T><<<grid_for(P.n_segs, blk), blk, 0, st>>>(P.n_segs, g, P.setup.ptr, P.segs.ptr, P.tile_off.ptr,
                                                            P.tile_cur.ptr, P.pairs.ptr, counters.ptr, P.n_tiles,
                                                            P.n_items, P.flags.ptr, P.slot_tiles.ptr);
    CBX_HIP(hipGetLastError());
    return CBX_OK;
}

static void resolve_opts(const cbx_draw_opts *o, int &tile, int &seg)
{
    // 64 x 64 tiles with a checkpoint every 8 columns: the fastest launch
    // shape measured on the 64-frame 2048 x 2048 workload
    tile = (o && o->tile) ? o->tile : 64;
    seg = (o && o->seg) ? o->seg : 8;
}

static int check_opts(int tile, int seg)
{
    if (!(tile == 32 || tile == 64)) return fail(CBX_EVALUE, "tile must be 32 or 64");
    if (!(seg == 8 || seg == 16)) return fail(CBX_EVALUE, "seg must be 8 or 16");
    return CBX_OK;
}

template <typename T, int K, int OV, int TILE, int SEG>
static void launch_tiles(T *out, const SegmentPlan<T> &P, T max_val, int threads, int chunk, bool fast,
                         hipStream_t st)
{
#define CBX_TILES(NT, CH, FST)                                                                                  \
    k_tiles<T, K, OV, TILE, SEG, NT, CH, FST><<<(unsigned)P.n_tiles, NT, 0, st>>>(                             \
        out, P.g, P.tile_off.ptr, P.pairs.ptr, P.segs.ptr, P.setup.ptr, max_val, P.flags.ptr)
    if constexpr (OV == 2) {
        if (threads == 64) CBX_TILES(64, 64, true);
        else if (threads == 128) CBX_TILES(128, 64, true);
        else if (chunk == 64) CBX_TILES(256, 64, true);
        else CBX_TILES(256, 256, true);
    } else if (fast) {
        if (threads == 64) CBX_TILES(64, 64, true);
        else if (threads == 128) CBX_TILES(128, 64, true);
        else if (chunk == 64) CBX_TILES(256, 64, true);
        else CBX_TILES(256, 256, true);
    } else {
        if (threads == 64) CBX_TILES(64, 64, false);
        else if (threads == 128) CBX_TILES(128, 64, false);
        else if (chunk == 64) CBX_TILES(256, 64, false);
        else CBX_TILES(256, 256, false);
    }
#undef CBX_TILES
}

template <typename T, int K, int OV>
static void launch_tiles_k(T *out, const SegmentPlan<T> &P, T max_val, int threads, int chunk, bool fast,
                           hipStream_t st)
{
    if (P.g.tile == 64 && P.g.seg == 16) launch_tiles<T, K, OV, 64, 16>(out, P, max_val, threads, chunk, fast, st);
    else if (P.g.tile == 64) launch_tiles<T, K, OV, 64, 8>(out, P, max_val, threads, chunk, fast, st);
    else if (P.g.seg == 16) launch_tiles<T, K, OV, 32, 16>(out, P, max_val, threads, chunk, fast, st);
    else launch_tiles<T, K, OV, 32, 8>(out, P, max_val, threads, chunk, fast, st);
}

template <typename T, int OV>
static void launch_tiles_ov(int kernel, T *out, const SegmentPlan<T> &P, T max_val, int threads, int chunk,
                            bool fast, hipStream_t st)
{
    switch (kernel) {
    case K_BIWEIGHT: launch_tiles_k<T, K_BIWEIGHT, OV>(out, P, max_val, threads, chunk, fast, st); break;
    case K_GAUSSIAN: launch_tiles_k<T, K_GAUSSIAN, OV>(out, P, max_val, threads, chunk, fast, st); break;
    case K_PARABOLIC: launch_tiles_k<T, K_PARABOLIC, OV>(out, P, max_val, threads, chunk, fast, st); break;
    case K_RECTANGULAR: launch_tiles_k<T, K_RECTANGULAR, OV>(out, P, max_val, threads, chunk, fast, st); break;
    case K_TRIANGULAR: launch_tiles_k<T, K_TRIANGULAR, OV>(out, P, max_val, threads, chunk, fast, st); break;
    default: launch_tiles_k<T, K_MIXED, OV>(out, P, max_val, threads, chunk, fast, st); break;
    }
}

template <typename T, bool WRITE>
static void launch_emit(int kernel, const SegmentPlan<T> &P, T max_val, int *cnt, const int64_t *off,
                        uint64_t *keys, T *vals, int32_t *lines_out, hipStream_t st)
{
    const int blk = 256;
    const int grid = grid_for(P.n_segs, blk);
#define CBX_EMIT(KK)                                                                                           \
    k_emit<T, KK, WRITE><<<grid, blk, 0, st>>>(P.segs.ptr, P.n_segs, P.setup.ptr, P.g.H, P.g.W, max_val, cnt, off, \
                                              keys, vals, lines_out, P.flags.ptr)
    switch (kernel) {
    case K_BIWEIGHT: CBX_EMIT(K_BIWEIGHT); break;
    case K_GAUSSIAN: CBX_EMIT(K_GAUSSIAN); break;
    case K_PARABOLIC: CBX_EMIT(K_PARABOLIC); break;
    case K_RECTANGULAR: CBX_EMIT(K_RECTANGULAR); break;
    default: CBX_EMIT(K_TRIANGULAR); break;
    }
#undef CBX_EMIT
}

// Raw pixel list: one entry per emitted pixel, in the reference's emission
// order (segments are stored line-major, in walk order).
template <typename T>
static int raw_list_pipeline(SegmentPlan<T> &P, int kernel, T max_val, int32_t *of, int32_t *oy, int32_t *ox, T *ov,
                             int32_t *ol, int64_t capacity, int64_t *count, hipStream_t st)
{
    *count = 0;
    if (P.n_segs == 0) return CBX_OK;
    const int blk = 256;
    DevBuf<int> cnt;
    DevBuf<int64_t> cnt64, off;
    CBX_HIP(cnt.alloc(P.n_segs + 1, st));
    CBX_HIP(cnt64.alloc(P.n_segs + 1, st));
    CBX_HIP(off.alloc(P.n_segs + 1, st));
    CBX_HIP(hipMemsetAsync(cnt.ptr + P.n_segs, 0, sizeof(int), st));
    launch_emit<T, false>(kernel, P, max_val, cnt.ptr, nullptr, nullptr, nullptr, nullptr, st);
    CBX_HIP(hipGetLastError());
    k_i32_to_i64<<<grid_for(P.n_segs + 1, blk), blk, 0, st>>>(cnt.ptr, cnt64.ptr, P.n_segs + 1);
    CBX_HIP(exclusive_scan_i64(cnt64.ptr, off.ptr, P.n_segs + 1, st));
    int64_t E = 0;
    int fl = 0;
    {
        const void *src[2] = {off.ptr + P.n_segs, P.flags.ptr};
        const int nb[2] = {8, 4};
        uint64_t hv[2];
        CBX_HIP(read_device_values(st, 2, src, nb, hv));
        E = (int64_t)hv[0];
        fl = (int)(uint32_t)hv[1];
    }
    int rc = flags_to_status(fl);
    if (rc) return rc;
    *count = E;
    if (E == 0 || E > capacity) return CBX_OK;
    DevBuf<uint64_t> keys;
    DevBuf<T> vals;
    DevBuf<int32_t> lof;
    CBX_HIP(keys.alloc(E, st));
    CBX_HIP(vals.alloc(E, st));
    const bool want_lines = (ol != nullptr);
    if (want_lines) CBX_HIP(lof.alloc(E, st));
    launch_emit<T, true>(kernel, P, max_val, nullptr, off.ptr, keys.ptr, vals.ptr, want_lines ? lof.ptr : nullptr,
                         st);
    CBX_HIP(hipGetLastError());
    k_raw_out<T><<<grid_for(E, blk), blk, 0, st>>>(keys.ptr, vals.ptr, lof.ptr, E, P.g.H, P.g.W, of, oy, ox, ov,
                                                  want_lines ? ol : nullptr);
    CBX_HIP(hipGetLastError());
    CBX_HIP(hipStreamSynchronize(st));
    return CBX_OK;
}

// tile_list = nullptr: one workgroup per plan tile; otherwise one per entry of
// tile_list (n_list tile ids).
template <typename T, int OV, int TILE, int SEG, bool LIST, int ACC = -1>
static void launch_ordered_k(int kernel, T *img, uint8_t *touched, int32_t *first, const SegmentPlan<T> &P,
                             T max_val, hipStream_t st, const int32_t *term_frame = nullptr,
                             const int *tile_list = nullptr, int64_t n_list = 0)
{
    const unsigned grid = (unsigned)(tile_list ? n_list : P.n_tiles);
#define CBX_ORD(KK)                                                                                             \
    k_tiles_ordered<T, KK, OV, TILE, SEG, LIST, ACC><<<grid, 64, 0, st>>>(                                      \
        img, touched, first, P.g, P.tile_off.ptr, P.items.ptr, P.segs.ptr, P.setup.ptr, max_val, P.flags.ptr,    \
        term_frame, tile_list)
    switch (kernel) {
    case K_BIWEIGHT: CBX_ORD(K_BIWEIGHT); break;
    case K_GAUSSIAN: CBX_ORD(K_GAUSSIAN); break;
    case K_PARABOLIC: CBX_ORD(K_PARABOLIC); break;
    case K_RECTANGULAR: CBX_ORD(K_RECTANGULAR); break;
    default: CBX_ORD(K_TRIANGULAR); break;
    }
#undef CBX_ORD
}

// Reference-order accumulation over a binned plan (tile 32, seg 16):
// image != nullptr accumulates into the caller's frames; otherwise a merged
// pixel list is compacted out of a dense scratch render.
template <typename T>
static int ordered_pipeline(SegmentPlan<T> &P, int kernel, int overlap, T max_val, T *image, int32_t *of, int32_t *oy,
                            int32_t *ox, T *ov, int32_t *ol, int64_t capacity, int64_t *count, hipStream_t st)
{
    if (count) *count = 0;
    if (P.n_segs == 0) return CBX_OK;
    k_sort_tile_items<<<(unsigned)P.n_tiles, 256, 0, st>>>(P.tile_off.ptr, P.items.ptr, P.flags.ptr, nullptr);
    CBX_HIP(hipGetLastError());
    const int64_t npix = P.g.n_frames * (int64_t)P.g.H * P.g.W;
    DevBuf<T> scratch;
    DevBuf<uint8_t> touched;
    DevBuf<int32_t> first;
    T *img = image;
    if (!image) {
        CBX_HIP(scratch.alloc(npix, st));
        CBX_HIP(touched.alloc(npix, st));
        if (ol) CBX_HIP(first.alloc(npix, st));
        // tiles without items never write their flags
        CBX_HIP(hipMemsetAsync(touched.ptr, 0, npix, st));
        img = scratch.ptr;
    }
    if (overlap == 0) {
        if (image) launch_ordered_k<T, 0, 32, 16, false>(kernel, img, nullptr, nullptr, P, max_val, st);
        else launch_ordered_k<T, 0, 32, 16, true>(kernel, img, touched.ptr, first.ptr, P, max_val, st);
    } else {
        if (image) launch_ordered_k<T, 1, 32, 16, false>(kernel, img, nullptr, nullptr, P, max_val, st);
        else launch_ordered_k<T, 1, 32, 16, true>(kernel, img, touched.ptr, first.ptr, P, max_val, st);
    }
    CBX_HIP(hipGetLastError());
    int fl = 0;
    {
        const void *src[1] = {P.flags.ptr};
        const int nb[1] = {4};
        uint64_t hv = 0;
        CBX_HIP(read_device_values(st, 1, src, nb, &hv));
        fl = (int)(uint32_t)hv;
    }
    if (fl & FLAG_TILE_FULL)
        return fail(CBX_ERUNTIME, "a tile holds more than 8192 segments; the ordered mode cannot sort it");
    int rc = flags_to_status(fl);
    if (rc || image) return rc;

    const int blk = 256;
    const int64_t rows = P.g.n_frames * (int64_t)P.g.H;
    DevBuf<int64_t> rcnt, roff;
    CBX_HIP(rcnt.alloc(rows + 1, st));
    CBX_HIP(roff.alloc(rows + 1, st));
    CBX_HIP(hipMemsetAsync(rcnt.ptr + rows, 0, sizeof(int64_t), st));
    k_row_count<<<grid_for(rows, blk), blk, 0, st>>>(touched.ptr, rows, P.g.W, rcnt.ptr);
    CBX_HIP(exclusive_scan_i64(rcnt.ptr, roff.ptr, rows + 1, st));
    int64_t U = 0;
    {
        const void *src[1] = {roff.ptr + rows};
        const int nb[1] = {8};
        uint64_t hv = 0;
        CBX_HIP(read_device_values(st, 1, src, nb, &hv));
        U = (int64_t)hv;
    }
    *count = U;
    if (U == 0 || U > capacity) return CBX_OK;
    k_row_write<T><<<grid_for(rows, blk), blk, 0, st>>>(touched.ptr, img, first.ptr, rows, P.g.H, P.g.W, roff.ptr, of,
                                                       oy, ox, ov, ol);
    CBX_HIP(hipGetLastError());
    CBX_HIP(hipStreamSynchronize(st));
    return CBX_OK;
}

// The tile kernel over a built unordered plan; kernel may be K_MIXED.
template <typename T>
static void run_tiles(T *out, const SegmentPlan<T> &P, int kernel, int overlap, T max_val, const cbx_draw_opts *opts,
                      hipStream_t st)
{
    const int threads = (opts && opts->threads) ? opts->threads : 256;
    const bool fast = !(opts && opts->exact_values);
    const int chunk = (opts && opts->chunk) ? opts->chunk : 64;
    // float sums on the fast path accumulate in fixed point, unless max_val
    // is not finite (inf * 0 must stay NaN, as in the reference)
    if constexpr (sizeof(T) == 4) {
        if (overlap == 0 && fast && std::isfinite(max_val)) {
            launch_tiles_ov<T, 2>(kernel, out, P, max_val, threads, chunk, fast, st);
            return;
        }
    }
    if (overlap == 0) launch_tiles_ov<T, 0>(kernel, out, P, max_val, threads, chunk, fast, st);
    else launch_tiles_ov<T, 1>(kernel, out, P, max_val, threads, chunk, fast, st);
}

template <typename T, typename I>
static int draw_lines_impl(T *out, int64_t n_frames, int64_t H, int64_t W, const T *lines, int64_t n_lines,
                           const I *idxs, T max_val, int kernel, int overlap, const cbx_draw_opts *opts, void *stream)
{
    int rc = check_common(n_frames, H, W, n_lines, idxs != nullptr, kernel, overlap);
    if (rc) return rc;
    int tile, seg;
    resolve_opts(opts, tile, seg);
    if ((rc = check_opts(tile, seg))) return rc;
    if (opts && !(opts->threads == 0 || opts->threads == 64 || opts->threads == 128 || opts->threads == 256))
        return fail(CBX_EVALUE, "threads must be 64, 128 or 256");
    if (n_lines == 0 || n_frames * H * W == 0) return CBX_OK;
    hipStream_t st = (hipStream_t)stream;
    const bool det = opts && opts->deterministic;
    const uint8_t *kernels = opts ? opts->kernels : nullptr;
    if (kernels && det) return fail(CBX_EVALUE, "per-line profiles are drawn by the tile path only (deterministic = 0)");
    if (det) {
        tile = 32;
        seg = 16;
    }
    SegmentPlan<T> P;
    // `out` may become ready on another stream: wait for it only before the
    // first kernel that touches it
    const hipEvent_t out_ready = opts ? (hipEvent_t)opts->out_ready : nullptr;
    if (det) {
        if ((rc = build_plan<T, I>(P, lines, idxs, kernel, n_lines, n_frames, H, W, tile, seg, true, st))) return rc;
        if (out_ready) CBX_HIP(hipStreamWaitEvent(st, out_ready, 0));
        if (P.n_segs == 0) return CBX_OK;
        return ordered_pipeline<T>(P, kernel, overlap, max_val, out, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                                   nullptr, st);
    }
    if ((rc = build_plan_unordered<T, I>(P, lines, idxs, kernels, kernel, n_lines, n_frames, H, W, tile, seg, st)))
        return rc;
    if (out_ready) CBX_HIP(hipStreamWaitEvent(st, out_ready, 0));
    if (P.n_segs == 0) return CBX_OK;
    {
        KernelTimer timer("draw_tiles", st);
        run_tiles<T>(out, P, kernels ? K_MIXED : kernel, overlap, max_val, opts, st);
    }
    CBX_HIP(hipGetLastError());
    return CBX_OK;
}

template <typename T, typename I>
static int draw_pixels_impl(int64_t n_frames, int64_t H, int64_t W, const T *lines, int64_t n_lines, const I *idxs,
                            T max_val, int kernel, int overlap, int merge, int32_t *of, int32_t *oy, int32_t *ox,
                            T *ov, int32_t *ol, int64_t capacity, int64_t *count, void *stream)
{
    if (!count) return fail(CBX_EVALUE, "count must not be NULL");
    *count = 0;
    int rc = check_common(n_frames, H, W, n_lines, idxs != nullptr, kernel, overlap);
    if (rc) return rc;
    if (n_lines == 0 || n_frames * H * W == 0) return CBX_OK;
    hipStream_t st = (hipStream_t)stream;
    SegmentPlan<T> P;
    if ((rc = build_plan<T, I>(P, lines, idxs, kernel, n_lines, n_frames, H, W, 32, 16, merge != 0, st))) return rc;
    if (merge) return ordered_pipeline<T>(P, kernel, overlap, max_val, nullptr, of, oy, ox, ov, ol, capacity, count, st);
    return raw_list_pipeline<T>(P, kernel, max_val, of, oy, ox, ov, ol, capacity, count, st);
}

// accumulate_lines, step 1: range checks of terms and frames (bad[0]: a term
// outside [0, n_terms), bad[1]: a frame outside [0, n_frames)), the frame of
// every term as int32 (frames32, n_terms entries) and, when comp != nullptr,
// the frame of every line, comp[i] = frames[terms[i]].
template <typename I>
__global__ void k_term_frames(const I *__restrict__ terms, const I *__restrict__ frames, int64_t n_lines,
                              int64_t n_terms, int64_t n_frames, I *__restrict__ comp,
                              int32_t *__restrict__ frames32, int *__restrict__ bad)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n_terms) {
        const I f = frames[i];
        const bool ok = f >= 0 && (int64_t)f < n_frames;
        if (!ok) atomicOr(bad + 1, 1);
        frames32[i] = ok ? (int32_t)f : 0;
    }
    if (i < n_lines) {
        const I t = terms[i];
        I f = 0;
        if (t < 0 || (int64_t)t >= n_terms) atomicOr(bad, 1);
        else f = frames[t];
        if (comp) comp[i] = (f < 0 || (int64_t)f >= n_frames) ? I(0) : f;
    }
}

// The ids of the plan tiles that hold items (in no fixed order), one atomic
// per workgroup of 256: list[0 .. *count).
__global__ void __launch_bounds__(256)
k_nonempty_tiles(const int *__restrict__ tile_off, int64_t n_tiles, int *__restrict__ list, int *__restrict__ count)
{
    __shared__ int wcnt[4];
    __shared__ int wbase[4];
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool ne = t < n_tiles && tile_off[t + 1] > tile_off[t];
    const uint64_t m = __ballot(ne);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wcnt[wave] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int total = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        const int base = total ? atomicAdd(count, total) : 0;
        wbase[0] = base;
        wbase[1] = base + wcnt[0];
        wbase[2] = base + wcnt[0] + wcnt[1];
        wbase[3] = base + wcnt[0] + wcnt[1] + wcnt[2];
    }
    __syncthreads();
    const int below = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    if (ne) list[wbase[wave] + below] = (int)t;
}

template <typename T, int IN_OV, int OUT_OV>
static void launch_accumulate(int kernel, T *out, const SegmentPlan<T> &P, T max_val, const int32_t *frames32,
                              const int *tile_list, int64_t n_list, hipStream_t st)
{
    launch_ordered_k<T, IN_OV, 32, 16, true, OUT_OV>(kernel, out, nullptr, nullptr, P, max_val, st, frames32,
                                                     tile_list, n_list);
}

// accumulate_lines (bresenham.cpp:151-256): lines grouped into terms, each
// term reduced over its lines (in_overlap), terms reduced into frames
// (out_overlap).  Equal overlaps compose to draw_lines with idxs =
// frames[terms] on the tile path.  Mixed overlaps bin the lines by (term,
// tile) and run the ordered tile kernel in accumulate mode: each term tile
// is reduced in LDS in reference line order from zero, like the reference's
// term buffer, and its touched pixels are combined into frame frames[term]
// with atomics (bresenham.cpp:230-248).  Terms reach a frame in no fixed
// order, as under the reference's OpenMP loop over terms.
// (term, tile) pairs of one mixed-overlap accumulate plan at most
constexpr int64_t MAX_TERM_TILES = int64_t(1) << 24;

// term of each line within the batch [t0, t1), or -1 (a line of another batch)
template <typename I>
__global__ void k_term_batch(const I *__restrict__ terms, int64_t n_lines, int64_t t0, int64_t t1,
                             I *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n_lines) return;
    const int64_t t = (int64_t)terms[i];
    out[i] = (t >= t0 && t < t1) ? (I)(t - t0) : (I)(-1);
}

template <typename T, typename I>
static int accumulate_batch(T *out, const T *lines, int64_t n_lines, const I *bidx, int64_t n_terms, bool skip,
                            const int32_t *frames32, int64_t H, int64_t W, T max_val, int kernel, int in_overlap,
                            hipStream_t st);

template <typename T, typename I>
static int accumulate_lines_impl(T *out, int64_t n_frames, int64_t H, int64_t W, const T *lines, int64_t n_lines,
                                 const I *terms, const I *frames, int64_t n_terms, T max_val, int kernel,
                                 int in_overlap, int out_overlap, const cbx_draw_opts *opts, void *stream)
{
    if (in_overlap < 0 || in_overlap > 1 || out_overlap < 0 || out_overlap > 1)
        return fail(CBX_EVALUE, "Invalid in_overlap and out_overlap keywords");
    if (n_terms < 0) return fail(CBX_EVALUE, "negative size");
    if (n_lines > 0 && !terms) return fail(CBX_EVALUE, "terms must not be NULL");
    if (n_terms > 0 && !frames) return fail(CBX_EVALUE, "frames must not be NULL");
    if (n_terms >= (int64_t(1) << 31)) return fail(CBX_EVALUE, "too many terms");
    int rc = check_common(n_frames, H, W, n_lines, true, kernel, in_overlap);
    if (rc) return rc;
    if (opts && opts->kernels) return fail(CBX_EVALUE, "accumulate_lines takes one profile per call");
    hipStream_t st = (hipStream_t)stream;
    const int64_t n_chk = n_lines > n_terms ? n_lines : n_terms;
    const bool same = in_overlap == out_overlap;
    DevBuf<I> comp;
    DevBuf<int32_t> frames32;
    if (n_chk > 0) {
        DevBuf<int> bad;
        CBX_HIP(bad.alloc(2, st));
        CBX_HIP(hipMemsetAsync(bad.ptr, 0, 2 * sizeof(int), st));
        if (same && n_lines > 0) CBX_HIP(comp.alloc(n_lines, st));
        CBX_HIP(frames32.alloc(n_terms > 0 ? n_terms : 1, st));
        k_term_frames<I><<<grid_for(n_chk, 256), 256, 0, st>>>(terms, frames, n_lines, n_terms, n_frames,
                                                                 comp.ptr, frames32.ptr, bad.ptr);
        CBX_HIP(hipGetLastError());
        const void *src[2] = {bad.ptr, bad.ptr + 1};
        const int nb[2] = {4, 4};
        uint64_t hv[2] = {0, 0};
        CBX_HIP(read_device_values(st, 2, src, nb, hv));
        if (hv[1]) return fail(CBX_EINDEX, "frames is outside of (0, n_frames)");
        if (hv[0]) return fail(CBX_EINDEX, "a term index is outside of (0, len(frames))");
    }
    if (n_lines == 0 || n_frames * H * W == 0) return CBX_OK;
    // max/max: the term buffer starts at zero and takes std::max(buffer, v),
    // so with max_val < 0 every drawn pixel holds 0 when its term ends: the
    // draw_lines max of zero values
    if (same) return draw_lines_impl<T, I>(out, n_frames, H, W, lines, n_lines, comp.ptr,
                                           (in_overlap == 1 && max_val < T(0)) ? T(0) : max_val, kernel,
                                           in_overlap, opts, stream);

    // The plan's frames are terms: (term, tile) lists scale with n_terms.
    // Above MAX_TERM_TILES pairs the terms go in batches, each with a plan of
    // its own (the lines of other terms skipped), in a nested workspace whose
    // memory the next batch reuses.
    const int64_t per_frame = ((H + 31) / 32) * ((W + 31) / 32);
    const int64_t batch = std::max<int64_t>(1, std::min<int64_t>(n_terms, MAX_TERM_TILES / per_frame));
    DevBuf<I> bterms;
    if (batch < n_terms) CBX_HIP(bterms.alloc(n_lines, st));
    for (int64_t t0 = 0; t0 < n_terms; t0 += batch) {
        const int64_t t1 = std::min(n_terms, t0 + batch);
        Workspace inner(st);
        const I *bidx = terms;
        if (batch < n_terms) {
            k_term_batch<I><<<grid_for(n_lines, 256), 256, 0, st>>>(terms, n_lines, t0, t1, bterms.ptr);
            CBX_HIP(hipGetLastError());
            bidx = bterms.ptr;
        }
        if ((rc = accumulate_batch<T, I>(out, lines, n_lines, bidx, t1 - t0, batch < n_terms, frames32.ptr + t0, H,
                                         W, max_val, kernel, in_overlap, st)))
            return rc;
    }
    return CBX_OK;
}

// One batch of terms of a mixed-overlap accumulate_lines: plan over the
// batch's (term, tile) pairs, tile sort, ordered render in accumulate mode.
// bidx: term of each line within the batch (negative: another batch's line).
template <typename T, typename I>
static int accumulate_batch(T *out, const T *lines, int64_t n_lines, const I *bidx, int64_t n_terms, bool skip,
                            const int32_t *frames32, int64_t H, int64_t W, T max_val, int kernel, int in_overlap,
                            hipStream_t st)
{
    int rc;
    SegmentPlan<T> P;
    if ((rc = build_plan<T, I>(P, lines, bidx, kernel, n_lines, n_terms, H, W, 32, 16, true, st, skip))) return rc;
    if (P.n_segs == 0) return CBX_OK;
    // (term, tile) pairs are mostly empty: launch over the non-empty ones only
    DevBuf<int> list, n_list;
    CBX_HIP(list.alloc(P.n_tiles, st));
    CBX_HIP(n_list.alloc(1, st));
    CBX_HIP(hipMemsetAsync(n_list.ptr, 0, sizeof(int), st));
    k_nonempty_tiles<<<grid_for(P.n_tiles, 256), 256, 0, st>>>(P.tile_off.ptr, P.n_tiles, list.ptr, n_list.ptr);
    CBX_HIP(hipGetLastError());
    int64_t nl = 0;
    {
        const void *src[1] = {n_list.ptr};
        const int nb[1] = {4};
        uint64_t hv = 0;
        CBX_HIP(read_device_values(st, 1, src, nb, &hv));
        nl = (int64_t)(uint32_t)hv;
    }
    if (nl < 0 || nl > P.n_tiles) return fail(CBX_ERUNTIME, "non-empty tile count out of range");
    if (nl == 0) return CBX_OK;
    k_sort_tile_items<<<(unsigned)nl, 256, 0, st>>>(P.tile_off.ptr, P.items.ptr, P.flags.ptr, list.ptr);
    CBX_HIP(hipGetLastError());
    {
        KernelTimer timer("accumulate_tiles", st);
        if (in_overlap == 0) launch_accumulate<T, 0, 1>(kernel, out, P, max_val, frames32, list.ptr, nl, st);
        else launch_accumulate<T, 1, 0>(kernel, out, P, max_val, frames32, list.ptr, nl, st);
    }
    CBX_HIP(hipGetLastError());
    int fl = 0;
    {
        const void *src[1] = {P.flags.ptr};
        const int nb[1] = {4};
        uint64_t hv = 0;
        CBX_HIP(read_device_values(st, 1, src, nb, &hv));
        fl = (int)(uint32_t)hv;
    }
    if (fl & FLAG_TILE_FULL)
        return fail(CBX_ERUNTIME, "a tile holds more than 8192 segments; accumulate_lines cannot sort it");
    return flags_to_status(fl);
}

template <typename T, int K>
static void launch_adjoint(const T *grad, const T *lines, const SegmentPlan<T> &P, const int *slot_base,
                           T max_val, double *partial, hipStream_t st)
{
    constexpr int NT = 256;
    if (P.g.tile == 64 && P.g.seg == 16)
        k_adjoint_tiles<T, K, 64, 16, NT, 16><<<(unsigned)P.n_tiles, NT, 0, st>>>(
            grad, lines, P.g, P.tile_off.ptr, P.pairs.ptr, P.segs.ptr, P.setup.ptr, slot_base, max_val, partial, P.flags.ptr);
    else if (P.g.tile == 64)
        k_adjoint_tiles<T, K, 64, 8, NT, 32><<<(unsigned)P.n_tiles, NT, 0, st>>>(
            grad, lines, P.g, P.tile_off.ptr, P.pairs.ptr, P.segs.ptr, P.setup.ptr, slot_base, max_val, partial, P.flags.ptr);
    else if (P.g.seg == 16)
        k_adjoint_tiles<T, K, 32, 16, NT, 16><<<(unsigned)P.n_tiles, NT, 0, st>>>(
            grad, lines, P.g, P.tile_off.ptr, P.pairs.ptr, P.segs.ptr, P.setup.ptr, slot_base, max_val, partial, P.flags.ptr);
    else
        k_adjoint_tiles<T, K, 32, 8, NT, 32><<<(unsigned)P.n_tiles, NT, 0, st>>>(
            grad, lines, P.g, P.tile_off.ptr, P.pairs.ptr, P.segs.ptr, P.setup.ptr, slot_base, max_val, partial, P.flags.ptr);
}

template <typename T, typename I>
static int adjoint_impl(const T *grad, int64_t n_frames, int64_t H, int64_t W, const T *lines, int64_t n_lines,
                        const I *idxs, T max_val, int kernel, T *grad_lines, const cbx_draw_opts *opts, void *stream)
{
    int rc = check_common(n_frames, H, W, n_lines, idxs != nullptr, kernel, 0);
    if (rc) return rc;
    int tile, seg;
    resolve_opts(opts, tile, seg);
    if ((rc = check_opts(tile, seg))) return rc;
    hipStream_t st = (hipStream_t)stream;
    if (n_lines == 0) return CBX_OK;
    CBX_HIP(hipMemsetAsync(grad_lines, 0, sizeof(T) * 5 * n_lines, st));
    if (n_frames * H * W == 0 || kernel == K_RECTANGULAR) return CBX_OK;
    SegmentPlan<T> P;
    DevBuf<int64_t> line_base;
    CBX_HIP(line_base.alloc(n_lines, st));
    if ((rc = build_plan_unordered<T, I>(P, lines, idxs, nullptr, kernel, n_lines, n_frames, H, W, tile, seg, st,
                                         line_base.ptr, true)))
        return rc;
    if (P.n_segs == 0) return CBX_OK;
    const int blk = 256;
    // each slot's first partial (tiles per slot from the binning), P.n_items
    // (< 2^31) in all
    DevBuf<int> sbase;
    CBX_HIP(sbase.alloc(P.n_segs + 1, st));
    CBX_HIP(exclusive_scan_total_i32(P.slot_tiles.ptr, nullptr, sbase.ptr, P.n_segs, st));
    // every (slot, tile) partial is written by exactly one tile
    DevBuf<double> partial;
    CBX_HIP(partial.alloc(5 * P.n_items, st));
    KernelTimer timer("adjoint_tiles", st);
    switch (kernel) {
    case K_BIWEIGHT: launch_adjoint<T, K_BIWEIGHT>(grad, lines, P, sbase.ptr, max_val, partial.ptr, st); break;
    case K_GAUSSIAN: launch_adjoint<T, K_GAUSSIAN>(grad, lines, P, sbase.ptr, max_val, partial.ptr, st); break;
    case K_PARABOLIC: launch_adjoint<T, K_PARABOLIC>(grad, lines, P, sbase.ptr, max_val, partial.ptr, st); break;
    default: launch_adjoint<T, K_TRIANGULAR>(grad, lines, P, sbase.ptr, max_val, partial.ptr, st); break;
    }
    CBX_HIP(hipGetLastError());
    k_line_grad<T><<<grid_for(n_lines * LG_LANES, blk), blk, 0, st>>>(line_base.ptr, P.setup.ptr, P.g, sbase.ptr,
                                                                      partial.ptr, n_lines, grad_lines, P.flags.ptr);
    CBX_HIP(hipGetLastError());
    int fl = 0;
    {
        const void *src[1] = {P.flags.ptr};
        const int nb[1] = {4};
        uint64_t hv = 0;
        CBX_HIP(read_device_values(st, 1, src, nb, &hv));
        fl = (int)(uint32_t)hv;
    }
    return flags_to_status(fl);
}

} // namespace cbx

// ---------------------------------------------------------------------------
// plans: the unordered tile path sized once, then drawn without host syncs,
// allocations or memsets
// ---------------------------------------------------------------------------
struct cbx_plan {
    cbx::SegmentPlan<float> P; // capacities in P.n_segs (slots) and P.n_items
    cbx::DevBuf<unsigned long long> counters;
    cbx::DevBuf<unsigned long long> scan_state; // the tile-offset scan's look-back state
    cbx::DevBuf<int64_t> line_base;             // each line's first slot in the last draw
    // the adjoint's buffers (allocated by the plan's first adjoint; the tiles
    // per slot, P.slot_tiles, are written by every draw)
    cbx::DevBuf<int> sbase;                     // each slot's first partial
    cbx::DevBuf<double> partial;                // five partial sums per (slot, tile)
    cbx::DevBuf<unsigned long long> adj_state;  // the slot-offset scan's look-back state
    int64_t n_lines_max = 0;
    int64_t bytes = 0;
    // the last draw: line count, profile (-1: per-line profiles)
    int64_t last_lines = -1;
    int last_kernel = -1;
    cbx_draw_opts opts{};
};

namespace cbx {

static int plan_create(cbx_plan **out_plan, int64_t n_frames, int64_t H, int64_t W, const float *lines,
                       int64_t n_lines, const int32_t *idxs, const uint8_t *kernels, int kernel,
                       const cbx_draw_opts *opts, double headroom, hipStream_t st)
{
    if (!out_plan) return fail(CBX_EVALUE, "plan must not be NULL");
    *out_plan = nullptr;
    int rc = check_common(n_frames, H, W, n_lines, idxs != nullptr, kernel, 0);
    if (rc) return rc;
    if (!(headroom >= 1.0) || headroom > 64.0) return fail(CBX_EVALUE, "headroom must be in [1, 64]");
    if (n_lines <= 0 || n_frames * H * W == 0) return fail(CBX_EVALUE, "a plan needs lines and a non-empty frame");
    int tile, seg;
    resolve_opts(opts, tile, seg);
    if ((rc = check_opts(tile, seg))) return rc;
    if (opts && !(opts->threads == 0 || opts->threads == 64 || opts->threads == 128 || opts->threads == 256))
        return fail(CBX_EVALUE, "threads must be 64, 128 or 256");
    if (opts && opts->deterministic) return fail(CBX_EVALUE, "plans draw with the tile path (deterministic = 0)");
    // sizing pass on the given lines (intermediates from the workspace cache)
    int64_t slots = 0, items = 0;
    {
        Workspace ws(st);
        SegmentPlan<float> S;
        if ((rc = build_plan_unordered<float, int32_t>(S, lines, idxs, kernels, kernel, n_lines, n_frames, H, W, tile,
                                                        seg, st)))
            return rc;
        slots = S.n_segs;
        items = S.n_items;
    }
    const int64_t slot_cap = std::max<int64_t>((int64_t)std::ceil((double)slots * headroom), slots + 1024);
    const int64_t item_cap = std::max<int64_t>((int64_t)std::ceil((double)items * headroom), items + 1024);
    if (slot_cap >= (int64_t(1) << 31) || item_cap >= (int64_t(1) << 31))
        return fail(CBX_EVALUE, "plan capacity exceeds 2^31 entries");
    cbx_plan *pl = new cbx_plan();
    SegmentPlan<float> &P = pl->P;
    Geometry &g = P.g;
    g.n_frames = n_frames;
    g.H = (int32_t)H;
    g.W = (int32_t)W;
    g.tile = tile;
    g.seg = seg;
    g.ntx = (int32_t)((W + tile - 1) / tile);
    g.nty = (int32_t)((H + tile - 1) / tile);
    P.n_lines = n_lines;
    P.n_tiles = n_frames * g.ntx * g.nty;
    const int64_t n_lists = P.n_tiles * LIST_SUBS;
    P.n_segs = slot_cap;
    P.n_items = item_cap;
    pl->n_lines_max = n_lines;
    if (opts) pl->opts = *opts;
    pl->opts.kernels = nullptr;
    pl->opts.tile = tile;
    pl->opts.seg = seg;
    hipError_t e = hipSuccess;
    auto take = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    // owned allocations (no workspace is open here)
    take(P.setup.alloc(n_lines, st));
    take(P.flags.alloc(1, st));
    take(P.segs.alloc(slot_cap, st));
    take(P.slot_tiles.alloc(slot_cap, st));
    take(P.tile_cnt.alloc(n_lists, st));
    take(P.tile_off.alloc(n_lists + 1, st));
    take(P.tile_cur.alloc(n_lists, st));
    take(P.pairs.alloc(item_cap, st));
    take(pl->counters.alloc(2, st));
    take(pl->line_base.alloc(n_lines, st));
    const size_t scan_words = scan_state_words(n_lists + 1);
    take(pl->scan_state.alloc(scan_words, st));
    if (e == hipSuccess) e = clear_scan_state(pl->scan_state.ptr, scan_words, st);
    if (e != hipSuccess) {
        delete pl;
        return fail(CBX_ERUNTIME, std::string("plan allocation: ") + hipGetErrorString(e));
    }
    pl->bytes = (int64_t)((sizeof(LineSetup<float>) + sizeof(int64_t)) * n_lines +
                          (sizeof(Segment<float>) + sizeof(int)) * slot_cap +
                          sizeof(int2) * item_cap + 3 * sizeof(int) * n_lists + sizeof(int) +
                          sizeof(unsigned long long) * scan_words);
    *out_plan = pl;
    return CBX_OK;
}

static int plan_draw(cbx_plan *pl, float *out, const float *lines, int64_t n_lines, const int32_t *idxs,
                     const uint8_t *kernels, float max_val, int kernel, int overlap, hipStream_t st)
{
    if (!pl) return fail(CBX_EVALUE, "plan must not be NULL");
    SegmentPlan<float> &P = pl->P;
    int rc = check_common(P.g.n_frames, P.g.H, P.g.W, n_lines, idxs != nullptr, kernel, overlap);
    if (rc) return rc;
    if (n_lines > pl->n_lines_max) return fail(CBX_EVALUE, "more lines than the plan was created for");
    if (n_lines == 0) return CBX_OK;
    const int blk = 256;
    const Geometry &g = P.g;
    const int64_t n_lists = P.n_tiles * LIST_SUBS;
    const int64_t reset_n = std::max<int64_t>(n_lists, (int64_t)pl->scan_state.n);
    k_plan_reset<<<grid_for(reset_n, blk), blk, 0, st>>>(n_lists, pl->counters.ptr, P.flags.ptr, P.tile_cnt.ptr,
                                                         P.tile_cur.ptr, pl->scan_state.ptr,
                                                         (int64_t)pl->scan_state.n);
    k_setup_lines<float, int32_t><<<grid_for(n_lines, blk), blk, 0, st>>>(lines, idxs, kernels, kernel, n_lines, g,
                                                                         P.setup.ptr, pl->counters.ptr, P.flags.ptr);
    k_walk_lines<float><<<grid_for(n_lines, blk), blk, 0, st>>>(n_lines, g, P.setup.ptr, P.segs.ptr, P.n_segs,
                                                                pl->counters.ptr, P.tile_cnt.ptr, P.flags.ptr,
                                                                pl->line_base.ptr);
    // the plan's own scan state (cleared by k_plan_reset above)
    CBX_HIP(scan_i32_with_state(P.tile_cnt.ptr, nullptr, P.tile_off.ptr, n_lists, n_lists + 1, pl->scan_state.ptr,
                                st));
    k_bin_slots<float><<<grid_for(P.n_segs, blk), blk, 0, st>>>(P.n_segs, g, P.setup.ptr, P.segs.ptr,
                                                                P.tile_off.ptr, P.tile_cur.ptr, P.pairs.ptr,
                                                                pl->counters.ptr, P.n_tiles, P.n_items, P.flags.ptr,
                                                                P.slot_tiles.ptr);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(st, &cap);
    if (cap == hipStreamCaptureStatusNone) {
        KernelTimer timer("draw_tiles", st);
        run_tiles<float>(out, P, kernels ? K_MIXED : kernel, overlap, max_val, &pl->opts, st);
    } else {
        run_tiles<float>(out, P, kernels ? K_MIXED : kernel, overlap, max_val, &pl->opts, st);
    }
    CBX_HIP(hipGetLastError());
    pl->last_lines = n_lines;
    pl->last_kernel = kernels ? -1 : kernel;
    return CBX_OK;
}

// The adjoint of the plan's last draw, on its segments and tile lists (no
// second walk): the same kernels as cbx_draw_lines_adjoint_f32, with no host
// synchronisation; errors of the draw leave grad_lines zero and are reported
// by plan_check.  The first call allocates the adjoint's buffers.
static int plan_adjoint(cbx_plan *pl, const float *grad, const float *lines, int64_t n_lines, const int32_t *idxs,
                        float max_val, int kernel, float *grad_lines, hipStream_t st)
{
    if (!pl) return fail(CBX_EVALUE, "plan must not be NULL");
    SegmentPlan<float> &P = pl->P;
    int rc = check_common(P.g.n_frames, P.g.H, P.g.W, n_lines, idxs != nullptr, kernel, 0);
    if (rc) return rc;
    if (n_lines == 0) return CBX_OK;
    if (pl->last_kernel < 0 && pl->last_lines >= 0)
        return fail(CBX_EVALUE, "the plan's last draw used per-line profiles: no adjoint");
    if (n_lines != pl->last_lines || kernel != pl->last_kernel)
        return fail(CBX_EVALUE, "the adjoint needs the line count and profile of the plan's last draw");
    if (!pl->partial.ptr) {
        hipError_t e = pl->sbase.alloc(P.n_segs + 1, st);
        if (e == hipSuccess) e = pl->partial.alloc(5 * P.n_items, st);
        if (e == hipSuccess) e = pl->adj_state.alloc(scan_state_words(P.n_segs + 1), st);
        if (e != hipSuccess) {
            pl->partial.release();
            return fail(CBX_ERUNTIME, std::string("plan adjoint allocation: ") + hipGetErrorString(e));
        }
        pl->bytes += (int64_t)(sizeof(int) * (P.n_segs + 1) + sizeof(double) * 5 * P.n_items +
                               sizeof(unsigned long long) * pl->adj_state.n);
    }
    {
        // the gradient and the scan's state, cleared by one kernel
        void *const zp[2] = {grad_lines, pl->adj_state.ptr};
        const int64_t zw[2] = {5 * n_lines, 2 * (int64_t)pl->adj_state.n};
        CBX_HIP(zero_words(st, 2, zp, zw));
    }
    if (kernel == K_RECTANGULAR) return CBX_OK;   // a flat profile: zero gradient
    const int blk = 256;
    CBX_HIP(scan_i32_with_state(P.slot_tiles.ptr, nullptr, pl->sbase.ptr, P.n_segs, P.n_segs + 1, pl->adj_state.ptr,
                                st));
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(st, &cap);
    {
        KernelTimer timer(cap == hipStreamCaptureStatusNone ? "adjoint_tiles" : nullptr, st);
        switch (kernel) {
        case K_BIWEIGHT: launch_adjoint<float, K_BIWEIGHT>(grad, lines, P, pl->sbase.ptr, max_val, pl->partial.ptr, st); break;
        case K_GAUSSIAN: launch_adjoint<float, K_GAUSSIAN>(grad, lines, P, pl->sbase.ptr, max_val, pl->partial.ptr, st); break;
        case K_PARABOLIC: launch_adjoint<float, K_PARABOLIC>(grad, lines, P, pl->sbase.ptr, max_val, pl->partial.ptr, st); break;
        default: launch_adjoint<float, K_TRIANGULAR>(grad, lines, P, pl->sbase.ptr, max_val, pl->partial.ptr, st); break;
        }
    }
    k_line_grad<float><<<grid_for(n_lines * LG_LANES, blk), blk, 0, st>>>(
        pl->line_base.ptr, P.setup.ptr, P.g, pl->sbase.ptr, pl->partial.ptr, n_lines, grad_lines, P.flags.ptr);
    CBX_HIP(hipGetLastError());
    return CBX_OK;
}

static int plan_check(cbx_plan *pl, hipStream_t st)
{
    if (!pl) return fail(CBX_EVALUE, "plan must not be NULL");
    int fl = 0;
    {
        const void *src[1] = {pl->P.flags.ptr};
        const int nb[1] = {4};
        uint64_t hv = 0;
        CBX_HIP(read_device_values(st, 1, src, nb, &hv));
        fl = (int)(uint32_t)hv;
    }
    return flags_to_status(fl);
}

} // namespace cbx

extern "C" {

int cbx_plan_create_f32(cbx_plan **plan, int64_t n_frames, int64_t height, int64_t width, const float *lines,
                        int64_t n_lines, const int32_t *idxs, const uint8_t *kernels, int kernel,
                        const cbx_draw_opts *opts, double headroom, void *stream)
{
    return cbx::plan_create(plan, n_frames, height, width, lines, n_lines, idxs, kernels, kernel, opts, headroom,
                            (hipStream_t)stream);
}

int cbx_plan_draw_f32(cbx_plan *plan, float *out, const float *lines, int64_t n_lines, const int32_t *idxs,
                      const uint8_t *kernels, float max_val, int kernel, int overlap, void *stream)
{
    return cbx::plan_draw(plan, out, lines, n_lines, idxs, kernels, max_val, kernel, overlap, (hipStream_t)stream);
}

int cbx_plan_adjoint_f32(cbx_plan *plan, const float *grad_out, const float *lines, int64_t n_lines,
                         const int32_t *idxs, float max_val, int kernel, float *grad_lines, void *stream)
{
    return cbx::plan_adjoint(plan, grad_out, lines, n_lines, idxs, max_val, kernel, grad_lines, (hipStream_t)stream);
}

int cbx_plan_check(cbx_plan *plan, void *stream) { return cbx::plan_check(plan, (hipStream_t)stream); }

int64_t cbx_plan_bytes(const cbx_plan *plan) { return plan ? plan->bytes : 0; }

void cbx_plan_destroy(cbx_plan *plan)
{
    if (!plan) return;
    (void)hipDeviceSynchronize();
    delete plan;
}

} // extern "C"

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int cbx_draw_lines_f32(float *out, int64_t n_frames, int64_t height, int64_t width, const float *lines,
                       int64_t n_lines, const int32_t *idxs, float max_val, int kernel, int overlap,
                       const cbx_draw_opts *opts, void *stream)
{
    cbx::Workspace ws((hipStream_t)stream);
    return cbx::draw_lines_impl<float, int32_t>(out, n_frames, height, width, lines, n_lines, idxs, max_val, kernel,
                                                overlap, opts, stream);
}

int cbx_draw_lines_f64(double *out, int64_t n_frames, int64_t height, int64_t width, const double *lines,
                       int64_t n_lines, const int64_t *idxs, double max_val, int kernel, int overlap,
                       const cbx_draw_opts *opts, void *stream)
{
    cbx::Workspace ws((hipStream_t)stream);
    return cbx::draw_lines_impl<double, int64_t>(out, n_frames, height, width, lines, n_lines, idxs, max_val, kernel,
                                                 overlap, opts, stream);
}

int cbx_draw_pixels_f32(int64_t n_frames, int64_t height, int64_t width, const float *lines, int64_t n_lines,
                        const int32_t *idxs, float max_val, int kernel, int overlap, int merge, int32_t *frame,
                        int32_t *y, int32_t *x, float *value, int32_t *line, int64_t capacity, int64_t *count,
                        void *stream)
{
    cbx::Workspace ws((hipStream_t)stream);
    return cbx::draw_pixels_impl<float, int32_t>(n_frames, height, width, lines, n_lines, idxs, max_val, kernel,
                                                 overlap, merge, frame, y, x, value, line, capacity, count, stream);
}

int cbx_draw_pixels_f64(int64_t n_frames, int64_t height, int64_t width, const double *lines, int64_t n_lines,
                        const int64_t *idxs, double max_val, int kernel, int overlap, int merge, int32_t *frame,
                        int32_t *y, int32_t *x, double *value, int32_t *line, int64_t capacity, int64_t *count,
                        void *stream)
{
    cbx::Workspace ws((hipStream_t)stream);
    return cbx::draw_pixels_impl<double, int64_t>(n_frames, height, width, lines, n_lines, idxs, max_val, kernel,
                                                  overlap, merge, frame, y, x, value, line, capacity, count, stream);
}

int cbx_accumulate_lines_f32(float *out, int64_t n_frames, int64_t height, int64_t width, const float *lines,
                             int64_t n_lines, const int32_t *terms, const int32_t *frames, int64_t n_terms,
                             float max_val, int kernel, int in_overlap, int out_overlap, const cbx_draw_opts *opts,
                             void *stream)
{
    cbx::Workspace ws((hipStream_t)stream);
    return cbx::accumulate_lines_impl<float, int32_t>(out, n_frames, height, width, lines, n_lines, terms, frames,
                                                      n_terms, max_val, kernel, in_overlap, out_overlap, opts,
                                                      stream);
}

int cbx_accumulate_lines_f64(double *out, int64_t n_frames, int64_t height, int64_t width, const double *lines,
                             int64_t n_lines, const int64_t *terms, const int64_t *frames, int64_t n_terms,
                             double max_val, int kernel, int in_overlap, int out_overlap, const cbx_draw_opts *opts,
                             void *stream)
{
    cbx::Workspace ws((hipStream_t)stream);
    return cbx::accumulate_lines_impl<double, int64_t>(out, n_frames, height, width, lines, n_lines, terms, frames,
                                                       n_terms, max_val, kernel, in_overlap, out_overlap, opts,
                                                       stream);
}

int cbx_draw_lines_adjoint_f32(const float *grad_out, int64_t n_frames, int64_t height, int64_t width,
                               const float *lines, int64_t n_lines, const int32_t *idxs, float max_val, int kernel,
                               float *grad_lines, const cbx_draw_opts *opts, void *stream)
{
    cbx::Workspace ws((hipStream_t)stream);
    return cbx::adjoint_impl<float, int32_t>(grad_out, n_frames, height, width, lines, n_lines, idxs, max_val, kernel,
                                             grad_lines, opts, stream);
}

int cbx_draw_lines_adjoint_f64(const double *grad_out, int64_t n_frames, int64_t height, int64_t width,
                               const double *lines, int64_t n_lines, const int64_t *idxs, double max_val,
                               int kernel, double *grad_lines, const cbx_draw_opts *opts, void *stream)
{
    cbx::Workspace ws((hipStream_t)stream);
    return cbx::adjoint_impl<double, int64_t>(grad_out, n_frames, height, width, lines, n_lines, idxs, max_val,
                                              kernel, grad_lines, opts, stream);
}

} // extern "C"
