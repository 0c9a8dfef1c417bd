// kernels.hip -- gfx950 kernels of the render path.
//
// Stage map (reference loop -> kernel):
//   vertex_shaders.hpp:16-71 (v_world, normal_world, frustum coords)  -> k_vertex
//   renderer.cpp:83-185 (visibility marks), binning counts          -> k_front
//   renderer.cpp:240-460 (clip, DDA setup) for binning              -> k_front / k_bin_fill
//   renderer.cpp:462-558 (spans, z test) + pixel shading            -> k_raster (per tile)
//   post_shaders.hpp:51-131 (depth of field)                        -> k_dof
// Every kernel works on a batch of jobs (one job = one viewport of one
// scene); a block finds its job by binary search over the job table, so the
// plain path is simply a batch of one.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "jobs.h"
#include "launch.h"

namespace sw {

// ------------------------------------------------------------ job lookup
enum { JV = 0, JT = 1, JD = 2, JTILE = 3 };
template <int F>
__device__ __forceinline__ uint32_t first_block(const JobDev &j)
{
    return F == JV ? j.vblock0 : F == JT ? j.tblock0 : F == JD ? j.dblock0 : j.tile_base;
}
// Every kernel takes the job table as a `const __restrict__` argument of its
// own as well: the table is not written while kernels run, and through such
// an argument the compiler reads uniform job fields with scalar loads.
template <int F>
__device__ __forceinline__ int find_job(const BatchDev &B, const JobDev *__restrict__ jobs, uint32_t blk)
{
    int lo = 0, hi = (int)B.n_jobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (first_block<F>(jobs[mid]) <= blk)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}
__device__ __forceinline__ int prim_of_vertex(const JobDev &J, uint32_t v)
{
    int lo = 0, hi = (int)J.n_prims - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (ro(J.prims)[mid].first_vertex <= v)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}
__device__ __forceinline__ int prim_of_tri(const JobDev &J, uint32_t t)
{
    int lo = 0, hi = (int)J.n_prims - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (ro(J.prims)[mid].first_tri <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ f3 ld3(const float4 *a, uint32_t i)
{
    const float4 v = ro4(a, i);
    return {v.x, v.y, v.z};
}

// Wave-aggregated atomicAdd.  Lanes holding the same key are grouped with
// ballots (ALU only); then every group leader issues its atomic in the same
// instruction, so the wave waits for one round trip instead of one per key.
// Grouping costs a lane exchange per distinct key, so it stops after
// SW_WAVE_ADD_GROUPS groups: the lanes left (many distinct keys, as a large
// triangle's tiles are) add 1 on their own in that same instruction.
// Returns this lane's slot (old value + rank among same-key lanes) when
// `want_slot`; inactive lanes get 0.
#ifndef SW_WAVE_ADD_GROUPS
#define SW_WAVE_ADD_GROUPS 4
#endif
template <bool want_slot>
__device__ __forceinline__ uint32_t wave_add(uint32_t *base, uint32_t key, bool active)
{
    const int lane = __lane_id();
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint64_t pending = __ballot(active);
    int my_leader = lane;
    uint32_t my_rank = 0, count = 1;   // a lane left ungrouped adds 1 itself
    for (int g = 0; pending && g < SW_WAVE_ADD_GROUPS; g++) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lkey = __shfl(key, leader);
        const uint64_t same = __ballot(active && key == lkey);
        if (active && key == lkey) {
            my_leader = leader;
            my_rank = (uint32_t)__popcll(same & below);
        }
        if (lane == leader)
            count = (uint32_t)__popcll(same);
        pending &= ~same;
    }
    const bool is_leader = active && my_leader == lane;
    if (!want_slot) {
        if (is_leader)
            atomicAdd(&base[key], count);
        return 0;
    }
    uint32_t old = 0;
    if (is_leader)
        old = atomicAdd(&base[key], count);
    old = __shfl(old, my_leader);
    return active ? old + my_rank : 0u;
}

__device__ __forceinline__ void stat_add(const BatchDev &B, int idx, unsigned long long v)
{
    unsigned long long *st = (unsigned long long *)(B.ctr + SW_CTR_STATS);   // 8-byte aligned
    atomicAdd(&st[(blockIdx.x & (SW_STAT_SLOTS - 1)) * 8 + idx], v);
}

// ------------------------------------------------------------- k_world
// original_to_world (vertex_shaders.hpp:16-33) and normal_world
// (:61-63).  Both depend only on the scene and its node poses, so they are
// computed when a scene is uploaded or re-posed, not per render.
__global__ __launch_bounds__(256) void k_world(WorldArgs W)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= W.n_vertices)
        return;
    int lo = 0, hi = (int)W.n_prims - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (W.prims[mid].first_vertex <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    const NodeDev &N = W.nodes[W.prims[lo].node];
    const f3 w = xform(N.M, f3{W.px[i], W.py[i], W.pz[i]});
    const f3 n = normal_world(N.RS, f3{W.nx[i], W.ny[i], W.nz[i]});
    // the texture coordinates ride in the fourth lanes, so a shading lane
    // gets a vertex's position, normal and coordinates in two loads
    W.world[i] = make_float4(w.x, w.y, w.z, W.tu[i]);
    W.nworld[i] = make_float4(n.x, n.y, n.z, W.tv[i]);
}

// ------------------------------------------------------------ k_vertex
// world_to_camera_or_frustum (vertex_shaders.hpp:35-52, 61-71) per viewport:
// frustum coordinates and cleared visibility marks.
// It also zeroes the tile counts and fill ranks and the counter words
// [w0, w0 + n_words) that the binning kernels after it accumulate into.
__global__ __launch_bounds__(256) void k_vertex(BatchDev B, uint32_t w0, uint32_t n_words, const JobDev *__restrict__ jobs)
{
    const uint32_t gi = blockIdx.x * 256u + threadIdx.x, stride = gridDim.x * 256u;
    for (uint32_t k = gi; k < B.n_tiles; k += stride) {
        B.tile_count[k * SW_TILE_PAD] = 0;
        B.tile_cursor[k * SW_TILE_PAD] = 0;
    }
    for (uint32_t k = gi; k < n_words; k += stride)
        B.ctr[w0 + k] = 0;
    const JobDev &J = jobs[find_job<JV>(B, jobs, blockIdx.x)];
    const uint32_t i = (blockIdx.x - J.vblock0) * 256u + threadIdx.x;
    if (i >= J.n_vertices)
        return;
    const float4 w4 = ro4(J.world, i);
    const f3 f = world_to_frustum(J.V, f3{w4.x, w4.y, w4.z});
    J.ndc[i] = make_float4(f.x, f.y, f.z, 0.0f);
    J.yes[i] = 0;
}

// A triangle's three vertices carry visibility marks (renderer.cpp:248-253
// draws a triangle only then).
__device__ __forceinline__ bool marked(const JobDev &J, uint32_t t)
{
    return ro(J.yes)[ro(J.tri)[3 * t]] && ro(J.yes)[ro(J.tri)[3 * t + 1]] && ro(J.yes)[ro(J.tri)[3 * t + 2]];
}

// ------------------------------------------------ fill_triangle on device
__device__ SVert orig_vertex(const JobDev &J, uint32_t i)
{
    SVert v;
    v.s = frustum_to_screen(J.V, ld3(J.ndc, i));
    v.a = v.b = i;
    v.cut = 0.0f;
    v.clip = 0;
    return v;
}
// new vertex between a and b (renderer.cpp:290-297 and world_to_viewport)
__device__ SVert clip_vertex(const JobDev &J, uint32_t a, uint32_t b, float cut)
{
    f3 wa = ld3(J.world, a), wb = ld3(J.world, b);
    f3 w = wa + (wb - wa) * cut;
    SVert v;
    v.s = frustum_to_screen(J.V, world_to_frustum(J.V, w));
    v.a = a;
    v.b = b;
    v.cut = cut;
    v.clip = 1;
    return v;
}

// fill_triangle (renderer.cpp:240-359): the sub-triangles handed to
// fill_triangle_2, in call order.  Returns how many (0, 1 or 2).
// BINNED: t came out of the binning, which lists a triangle only if all
// three of its vertices are marked (quick_extent, and this function with
// marks checked on the exact path), so the marks need no second look.
// The same with triangle t's vertex ids already loaded (i0, i1, i2).
template <bool BINNED = false>
__device__ int tri_subs_v(const JobDev &J, uint32_t i0, uint32_t i1, uint32_t i2, SubTri *out)
{
    if (!BINNED && !(ro(J.yes)[i0] && ro(J.yes)[i1] && ro(J.yes)[i2]))
        return 0;
    SVert v0 = orig_vertex(J, i0), v1 = orig_vertex(J, i1), v2 = orig_vertex(J, i2);
    bool front = cross_z_sign(v1.s - v0.s, v2.s - v0.s, -1);
    bool inv = false;
    if (v1.s.z > v0.s.z) { SVert x = v0; v0 = v1; v1 = x; inv = !inv; }
    if (v2.s.z > v1.s.z) { SVert x = v1; v1 = v2; v2 = x; inv = !inv; }
    if (v1.s.z > v0.s.z) { SVert x = v0; v0 = v1; v1 = x; inv = !inv; }
    if (v2.s.z >= SW_NEAR) {
        out[0].v[0] = v0;
        out[0].v[1] = v1;
        out[0].v[2] = v2;
        out[0].front = front;
        return 1;
    }
    if (v1.s.z < SW_NEAR) {
        float cut1 = (v0.s.z - 0.001f) / (v0.s.z - v1.s.z);
        SVert n1 = clip_vertex(J, v0.a, v1.a, cut1);
        float cut2 = (v0.s.z - 0.001f) / (v0.s.z - v2.s.z);
        SVert n2 = clip_vertex(J, v0.a, v2.a, cut2);
        bool f = cross_z_sign(n1.s - v0.s, n2.s - v0.s, -1);
        out[0].v[0] = v0;
        out[0].v[1] = n1;
        out[0].v[2] = n2;
        out[0].front = inv ? !f : f;
        return 1;
    }
    if (v2.s.z < SW_NEAR) {
        float cut0 = (v0.s.z - 0.001f) / (v0.s.z - v2.s.z);
        SVert n1 = clip_vertex(J, v0.a, v2.a, cut0);
        float cut1 = (v1.s.z - 0.001f) / (v1.s.z - v2.s.z);
        SVert n2 = clip_vertex(J, v1.a, v2.a, cut1);
        bool f0 = cross_z_sign(v1.s - v0.s, n2.s - v0.s, -1);
        out[0].v[0] = v0;
        out[0].v[1] = v1;
        out[0].v[2] = n2;
        out[0].front = inv ? !f0 : f0;
        bool f1 = cross_z_sign(n2.s - v0.s, n1.s - v0.s, -1);
        out[1].v[0] = v0;
        out[1].v[1] = n2;
        out[1].v[2] = n1;
        out[1].front = inv ? !f1 : f1;
        return 2;
    }
    return 0;   // NaN depth: no branch of the reference is taken
}
template <bool BINNED = false>
__device__ __forceinline__ int tri_subs(const JobDev &J, uint32_t t, SubTri *out)
{
    return tri_subs_v<BINNED>(J, ro(J.tri)[3 * t], ro(J.tri)[3 * t + 1], ro(J.tri)[3 * t + 2], out);
}

// Rows a sub-triangle's halves process, and a conservative column range.
struct Extent {
    int y0, y1;   // [y0, y1)
    int x0, x1;   // [x0, x1]
};
// Columns of a triangle's tile rectangle: the vertex x range widened by
// SW_RECT_MARGIN pixels plus a bound of the DDA's accumulated rounding
// ((rows + 8) add errors of at most 2^-20 |x|max each, 16 times the float
// half-ulp).  A covered pixel x satisfies ceil(left.x) <= x < ceil(right.x)
// with both sides within the vertex x range up to that rounding, so the
// bound holds with no margin at all; the margin is spare.  A brute-force
// check against the rasteriser on adversarial triangles pins it.
#define SW_RECT_MARGIN 0.5f

__device__ bool sub_extent(const Setup &S, const ViewXf &V, Extent &E)
{
    int ya = INT32_MAX, yb = INT32_MIN;
    if (S.has_up && S.up_y1 > S.up_y0) {
        ya = min(ya, S.up_y0);
        yb = max(yb, S.up_y1);
    }
    if (S.has_lo && S.lo_y1 > S.lo_y0) {
        ya = min(ya, S.lo_y0);
        yb = max(yb, S.lo_y1);
    }
    if (ya >= yb)
        return false;
    E.y0 = ya;
    E.y1 = yb;
    float xmin = fminf(S.v[0].s.x, fminf(S.v[1].s.x, S.v[2].s.x));
    float xmax = fmaxf(S.v[0].s.x, fmaxf(S.v[1].s.x, S.v[2].s.x));
    // The DDA reaches row k after k float adds; each add is off by at most
    // half an ulp of the largest |x| on the walk.  Widen by that bound.
    float X = fmaxf(fmaxf(fabsf(xmin), fabsf(xmax)), 1.0f);
    float slack = SW_RECT_MARGIN + (float)(yb - ya + 8) * X * (1.0f / 1048576.0f);
    if (!(xmin == xmin) || !(xmax == xmax) || X > 1e30f) {
        E.x0 = V.vx;
        E.x1 = V.vx + V.vw - 1;
        return true;
    }
    float fx0 = floorf(xmin - slack), fx1 = ceilf(xmax + slack);
    E.x0 = fx0 < (float)V.vx ? V.vx : (int)fx0;
    E.x1 = fx1 > (float)(V.vx + V.vw - 1) ? V.vx + V.vw - 1 : (int)fx1;
    return E.x0 <= E.x1;
}

#define RECT_EMPTY 0xFFFFFFFFu

// ------------------------------------------------------ tile rectangles
// Tile rectangle of a triangle, conservative but never too small: rows are
// exact (ceil of vertex y, clipped to the viewport), columns are the vertex
// x range widened by the DDA's worst-case accumulated rounding.
// Near-clipped or non-finite triangles take the exact fill_triangle path.
// The extent of an unclipped, finite triangle from its screen points; sets
// `exact` when the triangle needs fill_triangle's exact path (near-clipped
// or non-finite).  `drawn`: fill_triangle_2 would draw it (y0 != y2).
__device__ bool quick_extent(const JobDev &J, f3 a, f3 b, f3 c, Extent &E, uint32_t &drawn, bool &exact)
{
    exact = false;
    drawn = 0;
    const float big = 1073741824.0f;
    bool finite = fabsf(a.x) < big && fabsf(b.x) < big && fabsf(c.x) < big && fabsf(a.y) < big &&
                  fabsf(b.y) < big && fabsf(c.y) < big;
    if (!finite || !(fminf(a.z, fminf(b.z, c.z)) >= SW_NEAR) || a.z != a.z || b.z != b.z || c.z != c.z) {
        exact = true;
        return false;
    }
    const float ymin = fminf(a.y, fminf(b.y, c.y)), ymax = fmaxf(a.y, fmaxf(b.y, c.y));
    const int iy0 = cvtt(ceilf(ymin)), iy2 = cvtt(ceilf(ymax));
    if (iy0 == iy2)
        return false;
    drawn = 1;
    const int y0 = max(iy0, J.V.vy), y1 = min(iy2, J.V.vy + J.V.vh);
    if (y0 >= y1)
        return false;
    E.y0 = y0;
    E.y1 = y1;
    const float xmin = fminf(a.x, fminf(b.x, c.x)), xmax = fmaxf(a.x, fmaxf(b.x, c.x));
    const float X = fmaxf(fmaxf(fabsf(xmin), fabsf(xmax)), 1.0f);
    const float slack = SW_RECT_MARGIN + (float)(iy2 - iy0 + 8) * X * (1.0f / 1048576.0f);
    const float fx0 = floorf(xmin - slack), fx1 = ceilf(xmax + slack);
    E.x0 = fx0 < (float)J.V.vx ? J.V.vx : (int)fx0;
    E.x1 = fx1 > (float)(J.V.vx + J.V.vw - 1) ? J.V.vx + J.V.vw - 1 : (int)fx1;
    return E.x0 <= E.x1;
}

// The exact path: the union extent of fill_triangle's sub-triangles and how
// many of them fill_triangle_2 draws.  Marks are not looked at here.
__device__ bool exact_extent(const JobDev &J, uint32_t t, Extent &E, uint32_t &drawn)
{
    SubTri st[2];
    const int ns = tri_subs<true>(J, t, st);
    int ya = INT32_MAX, yb = INT32_MIN, xa = INT32_MAX, xb = INT32_MIN;
    drawn = 0;
#pragma unroll
    for (int s = 0; s < 2; s++) {
        Setup S;
        if (s >= ns || !setup_subtri(st[s], J.V, S))
            continue;
        drawn++;
        Extent e;
        if (!sub_extent(S, J.V, e))
            continue;
        ya = min(ya, e.y0);
        yb = max(yb, e.y1);
        xa = min(xa, e.x0);
        xb = max(xb, e.x1);
    }
    E.y0 = ya;
    E.y1 = yb;
    E.x0 = xa;
    E.x1 = xb;
    return ya < yb && xa <= xb;
}

// The tile rows of the k_front/k_bin_fill rectangle, shared by both.
__device__ __forceinline__ uint2 tile_rect(const JobDev &J, const Extent &E, int &tx0, int &tx1, int &ty0, int &ty1)
{
    // tile edges are powers of two and the offsets are not negative: shifts
    const int sx = __builtin_ctz((unsigned)J.tw), sy = __builtin_ctz((unsigned)J.th);
    tx0 = (E.x0 - J.V.vx) >> sx;
    tx1 = (E.x1 - J.V.vx) >> sx;
    ty0 = (E.y0 - J.V.vy) >> sy;
    ty1 = (E.y1 - 1 - J.V.vy) >> sy;
    return make_uint2((uint32_t)tx0 | ((uint32_t)ty0 << 16), (uint32_t)tx1 | ((uint32_t)ty1 << 16));
}

// ------------------------------------------------------ tile count scan
// Exclusive scan of the tile counts by one block, in passes of NT x
// SCAN_PER tiles: each thread loads its SCAN_PER consecutive counts at once
// (one memory round trip per pass), the block scans the per-thread sums
// (shuffles within each wave, then the wave totals), and each thread writes
// its tiles' offsets (densely: the fill ranks, zeroed by k_vertex, sit in
// the padded lines and are added to the offsets by k_bin_fill).
#define SCAN_PER 16
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d);
        if (lane >= d)
            v += o;
    }
    return v;
}

// Claim order of the raster's tiles: largest bin counts first (a counting
// sort on a logarithmic bucket of the count, four buckets per power of
// two), so that the costliest tiles start first and the last ones to start
// are cheap -- the persistent raster's makespan then comes close to its
// mean load instead of the heaviest tile started last.
#define SW_ORDER_BUCKETS 128
static_assert(SW_ORDER_BUCKETS == 2 * 64, "the bucket scan takes two buckets per lane of one wave");
__device__ __forceinline__ uint32_t count_bucket(uint32_t c)
{
    if (c < 4)
        return c;
    const uint32_t e = 31u - (uint32_t)__clz(c);   // >= 2
    return 4u * (e - 1u) + ((c >> (e - 2u)) & 3u);   // <= 123
}

// SCAN_PER consecutive tile counts from tile i0 on (zero past n), read with
// device-scope atomic loads (the counts were added at the memory side in
// this same kernel: an XCD's L2 may hold older lines), all issued before
// the first is used.
__device__ __forceinline__ void load_counts(const BatchDev &B, uint32_t i0, uint32_t n, uint32_t (&v)[SCAN_PER])
{
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++)
        v[k] = i0 + k < n ? __hip_atomic_load(&B.tile_count[(i0 + k) * SW_TILE_PAD], __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT)
                          : 0u;
}

// The exclusive scan of the tile counts into the bin offsets, by one block
// of NT threads.
template <int NT>
__device__ void scan_offsets(const BatchDev &B)
{
    constexpr int NW = NT / 64;
    __shared__ uint32_t wsum[NW];
    __shared__ uint32_t s_carry;
    const uint32_t n = B.n_tiles;
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    if (threadIdx.x == 0)
        s_carry = 0;
    __syncthreads();
    uint32_t v[SCAN_PER];
    for (uint32_t base = 0; base < n; base += (uint32_t)NT * SCAN_PER) {
        const uint32_t i0 = base + threadIdx.x * SCAN_PER;
        load_counts(B, i0, n, v);
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < SCAN_PER; k++)
            sum += v[k];
        const uint32_t incl = wave_incl_scan(sum);
        if (lane == 63)
            wsum[wave] = incl;
        __syncthreads();
        if (wave == 0) {
            const uint32_t w = lane < NW ? wsum[lane] : 0u;
            const uint32_t wi = wave_incl_scan(w);
            if (lane < NW)
                wsum[lane] = wi - w;   // exclusive prefix of the wave totals
        }
        __syncthreads();
        const uint32_t carry = s_carry;
        uint32_t run = carry + wsum[wave] + incl - sum;
#pragma unroll
        for (int k = 0; k < SCAN_PER; k++) {
            if (i0 + k < n) {
                B.tile_offset[i0 + k] = run;
                run += v[k];
            }
        }
        __syncthreads();   // everyone has read s_carry and wsum
        if (threadIdx.x == NT - 1)
            s_carry = run;   // the last thread's running sum ends the pass
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        B.tile_offset[n] = s_carry;
        B.ctr[SW_CTR_TOTAL] = s_carry;   // total bin entries
    }
}

// The raster's claim order, by one block of NT threads: the tiles sorted by
// a log bucket of their bin counts, the largest bucket first (order within
// a bucket is arbitrary: the raster's result does not depend on the claim
// order).
template <int NT>
__device__ void order_tiles(const BatchDev &B)
{
    __shared__ uint32_t s_hist[SW_ORDER_BUCKETS];
    const uint32_t n = B.n_tiles;
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    for (int b = threadIdx.x; b < SW_ORDER_BUCKETS; b += NT)
        s_hist[b] = 0;
    __syncthreads();
    uint32_t v[SCAN_PER];
    for (uint32_t base = 0; base < n; base += (uint32_t)NT * SCAN_PER) {
        const uint32_t i0 = base + threadIdx.x * SCAN_PER;
        load_counts(B, i0, n, v);
#pragma unroll
        for (int k = 0; k < SCAN_PER; k++)
            if (i0 + k < n)
                atomicAdd(&s_hist[count_bucket(v[k])], 1u);
    }
    __syncthreads();
    // first place of each bucket, the largest bucket first: wave 0 scans the
    // buckets in descending order, two per lane
    if (wave == 0) {
        const int b0 = SW_ORDER_BUCKETS - 1 - 2 * lane, b1 = b0 - 1;
        const uint32_t h0 = s_hist[b0], h1 = s_hist[b1];
        const uint32_t incl = wave_incl_scan(h0 + h1);
        s_hist[b0] = incl - h0 - h1;
        s_hist[b1] = incl - h1;
    }
    __syncthreads();
    // every tile takes the next place of its bucket; with one pass the
    // counts are still in registers
    for (uint32_t base = 0; base < n; base += (uint32_t)NT * SCAN_PER) {
        const uint32_t i0 = base + threadIdx.x * SCAN_PER;
        if (n > (uint32_t)NT * SCAN_PER)
            load_counts(B, i0, n, v);
#pragma unroll
        for (int k = 0; k < SCAN_PER; k++)
            if (i0 + k < n)
                B.tile_order[atomicAdd(&s_hist[count_bucket(v[k])], 1u)] = i0 + k;
    }
}

// ------------------------------------------------------------- k_front
// Per triangle, in one pass: the visibility marks (renderer.cpp:83-185: a
// triangle passing the frustum test and double-sided or front-facing marks
// its three vertices), its tile rectangle (conservative but never too small:
// rows exact from ceil of vertex y, clipped to the viewport; columns the
// vertex x range widened by the DDA's worst-case accumulated rounding;
// near-clipped or non-finite triangles take the exact fill_triangle path)
// and one bin count per covered tile.  Whether a triangle is drawn depends
// on marks other triangles set in this same pass, so every triangle with a
// rectangle is counted here; k_bin_fill, which runs after all marks are
// set, places only triangles whose three vertices are marked, and the
// raster checks the marks of big-list triangles.  The last block to finish
// scans the tile counts into the bin offsets (the claim order is written by
// k_bin_fill's last block, off this kernel's critical path).
__global__ __launch_bounds__(256) void k_front(BatchDev B, const JobDev *__restrict__ jobs)
{
    const int jidx = find_job<JT>(B, jobs, blockIdx.x);
    const JobDev &J = jobs[jidx];
    const uint32_t t = (blockIdx.x - J.tblock0) * 256u + threadIdx.x;
    uint32_t big = 0;
    int ntiles = 0, tx0 = 0, tx1 = 0, ty0 = 0;
    bool self_marked = false;   // marked its own vertices: drawn whatever other triangles do
    if (t < J.n_tris) {
        const uint32_t i0 = ro(J.tri)[3 * t], i1 = ro(J.tri)[3 * t + 1], i2 = ro(J.tri)[3 * t + 2];
        const f3 a = ld3(J.ndc, i0), b = ld3(J.ndc, i1), c = ld3(J.ndc, i2);
        if (inside_frustum(a, b, c) &&
            (ro(J.prims)[prim_of_tri(J, t)].double_sided || cross_z_sign(b - a, c - a, 1))) {
            J.yes[i0] = 1;
            J.yes[i1] = 1;
            J.yes[i2] = 1;
            self_marked = true;
        }
        Extent E;
        bool exact;
        uint32_t drawn;
        bool have = quick_extent(J, frustum_to_screen(J.V, a), frustum_to_screen(J.V, b),
                                 frustum_to_screen(J.V, c), E, drawn, exact);
        if (exact)
            have = exact_extent(J, t, E, drawn);
        uint2 r = make_uint2(RECT_EMPTY, 0);
        if (have) {
            int ty1;
            r = tile_rect(J, E, tx0, tx1, ty0, ty1);
            ntiles = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
            if (ntiles > (int)B.big_tiles) {
                ntiles = 0;
                big = 1;
            }
        }
        J.rect[t] = r;
    }
    // big-list append, one atomic per wave (all lanes of a block share the job)
    {
        const uint32_t pos = wave_add<true>(B.ctr + SW_CTR_BIG0, (uint32_t)jidx, big != 0);
        if (big)
            J.big[pos] = t;
    }
    // one aggregated atomic per distinct tile per wave and round; with a
    // tile capacity the count's old value places the entry
    bool overflow = false;
    for (int k = 0; __any(k < ntiles); k++) {
        const bool act = k < ntiles;
        uint32_t key = 0;
        if (act) {
            const int w = tx1 - tx0 + 1;
            const int ky = k == 0 ? 0 : k / w, kx = k == 0 ? 0 : k % w;   // k is the same in every lane
            key = J.tile_base + (ty0 + ky) * J.tiles_x + (tx0 + kx);
        }
        if (B.fixed) {
            const uint32_t rank = wave_add<true>(B.tile_count, key * SW_TILE_PAD, act);
            if (act && rank < J.tile_cap)   // top bit: the raster must look at the marks
                B.bins_fixed[J.bin_base + (size_t)(key - J.tile_base) * J.tile_cap + rank] =
                    self_marked ? t : t | SW_BIN_CHECK;
            overflow |= act && rank >= J.tile_cap;
        } else {
            wave_add<false>(B.tile_count, key * SW_TILE_PAD, act);
        }
    }
    if (__any(overflow) && __lane_id() == 0) {   // every lane of a block has the block's job
        atomicOr(B.ctr + SW_CTR_BIG0 + B.n_jobs + jidx, 1u);
        atomicOr(B.ctr + SW_CTR_OVERFLOW, 1u);
    }
    // The last block to finish scans.  The counts are device-scope atomics,
    // performed at the memory side, not in an XCD's L2: each wave waits
    // for its own to complete before the block takes its ticket, and the
    // scan reads the counts with device-scope atomic loads.  (A device-scope
    // fence here would write back and invalidate every XCD's L2 per block.)
    // Tickets are taken in two levels (the blocks of group k = index mod
    // SW_TICKET_GROUPS, then the groups), so that no counter sees more than
    // about gridDim / SW_TICKET_GROUPS atomics.
    __shared__ uint32_t s_last;
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t G = gridDim.x, k = blockIdx.x % SW_TICKET_GROUPS;
        const uint32_t in_group = (G - 1u - k) / SW_TICKET_GROUPS + 1u;   // blocks k, k + 32, ... below G
        const uint32_t groups = G < SW_TICKET_GROUPS ? G : SW_TICKET_GROUPS;
        uint32_t last = 0;
        if (atomicAdd(B.ctr + SW_CTR_TICKETS + 32 * k, 1u) == in_group - 1u)
            last = atomicAdd(B.ctr + SW_CTR_TICKETS + 32 * SW_TICKET_GROUPS, 1u) == groups - 1u ? 1u : 0u;
        s_last = last;
    }
    __syncthreads();
    if (!s_last)
        return;
    // the offsets serve the filled layout only
    if (!B.fixed || __hip_atomic_load(B.ctr + SW_CTR_OVERFLOW, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        scan_offsets<256>(B);
}

// ----------------------------------------------------------- k_bin_fill
// Block 0 writes the raster's claim order instead (the counts are final
// once k_front has ended): dispatched first, it runs beside the fill, not
// after it.  Block b > 0 fills for triangle block b - 1.
__global__ __launch_bounds__(256) void k_bin_fill(BatchDev B, const JobDev *__restrict__ jobs)
{
    if (blockIdx.x == 0) {
        order_tiles<256>(B);
        return;
    }
    // k_front placed every entry of a job unless one of its tiles
    // overflowed: then only the statistics, if asked for, are counted here
    const uint32_t blk = blockIdx.x - 1;
    const int jidx = find_job<JT>(B, jobs, blk);
    const bool place = !B.fixed || B.ctr[SW_CTR_BIG0 + B.n_jobs + jidx] != 0;
    if (!place && !B.stats)
        return;
    const JobDev &J = jobs[jidx];
    const uint32_t t = (blk - J.tblock0) * 256u + threadIdx.x;
    int ntiles = 0, tx0 = 0, tx1 = 0, ty0 = 0;
    uint32_t drawn = 0, big = 0;
    if (t < J.n_tris) {
        const uint2 r = ro2(J.rect, t);
        const bool m = (r.x != RECT_EMPTY || B.stats) && marked(J, t);   // every mark is set by now
        if (r.x != RECT_EMPTY && m) {
            tx0 = r.x & 0xFFFF;
            ty0 = r.x >> 16;
            tx1 = r.y & 0xFFFF;
            const int ty1 = r.y >> 16;
            ntiles = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
            if (ntiles > (int)B.big_tiles) {
                ntiles = 0;
                big = 1;
            }
        }
        if (B.stats && m) {
            // drawn sub-triangles (fill_triangle_2 calls that find a scanline)
            const uint32_t i0 = ro(J.tri)[3 * t], i1 = ro(J.tri)[3 * t + 1], i2 = ro(J.tri)[3 * t + 2];
            Extent E;
            bool exact;
            quick_extent(J, frustum_to_screen(J.V, ld3(J.ndc, i0)), frustum_to_screen(J.V, ld3(J.ndc, i1)),
                         frustum_to_screen(J.V, ld3(J.ndc, i2)), E, drawn, exact);
            if (exact)
                exact_extent(J, t, E, drawn);
        }
    }
    uint32_t entries = 0;
    for (int k = 0; __any(k < ntiles); k++) {
        const bool act = k < ntiles;
        if (!place) {   // (wave-uniform)
            entries += act ? 1u : 0u;
            continue;
        }
        uint32_t key = 0;
        if (act) {
            const int w = tx1 - tx0 + 1;
            const int ky = k == 0 ? 0 : k / w, kx = k == 0 ? 0 : k % w;   // k is the same in every lane
            key = J.tile_base + (ty0 + ky) * J.tiles_x + (tx0 + kx);
            entries++;
        }
        // the entry's place: the tile's offset (read while the rank's atomic is
        // in flight) plus its rank among the tile's entries so far
        const uint32_t off = act ? ro(B.tile_offset)[key] : 0u;
        const uint32_t rank = wave_add<true>(B.tile_cursor, key * SW_TILE_PAD, act);
        const uint32_t pos = off + rank;
        if (act && pos < B.bin_capacity)
            B.bins[pos] = t;
    }
    if (!B.stats)
        return;
    __shared__ uint32_t s_acc[3];
    if (threadIdx.x < 3)
        s_acc[threadIdx.x] = 0;
    __syncthreads();
    if (drawn)
        atomicAdd(&s_acc[0], drawn);
    if (entries)
        atomicAdd(&s_acc[1], entries);
    if (big)
        atomicAdd(&s_acc[2], big);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_acc[0])
            stat_add(B, SW_ST_DRAWN, s_acc[0]);
        if (s_acc[1])
            stat_add(B, SW_ST_BINS, s_acc[1]);
        if (s_acc[2])
            stat_add(B, SW_ST_BIG, s_acc[2]);
    }
}

// `by` divides by fdepth - 1 (the same for the whole job); `fast` when
// 1 <= fdepth - 1 <= 2^20, where FDiv's quotient equals '/' (the numerator
// t - 1 lies in [2^-23, fdepth - 1] on that branch).
__device__ __forceinline__ float blur_factor(float z, float fdist, float fdepth, float rmax, const FDiv &by,
                                             bool fast)
{
    float t = fabsf(fdist - z);
    // inv_lerp_clipped(1, fdepth, t)
    float x;
    if (1.0f == fdepth)
        x = 0.5f;
    else if (t <= 1.0f)
        x = 0.0f;
    else if (t >= fdepth)
        x = 1.0f;
    else
        x = fast ? by(t - 1.0f) : (t - 1.0f) / (fdepth - 1.0f);
    // lerp_clipped(0, rmax, x)
    if (x <= 0)
        return 0.0f;
    if (x >= rmax)
        return rmax;
    return 0.0f + rmax * x;
}

// ------------------------------------------------------------- k_raster
// One workgroup per screen tile.  Phase 1 replays every binned sub-triangle's
// DDA over the tile's rows and columns (exact float sequence) and keeps the
// minimum (z, draw sequence) per pixel with an LDS 64-bit atomic min.
// Phase 2 shades each pixel's winner at its own progress and writes colour,
// depth (and the winner's sequence) once, coalesced.
struct TileRect { int x0, y0, x1, y1; };   // [x0,x1) x [y0,y1), absolute pixels

// Rows [wy0, wy1) of the tile only, so that one large triangle can be split
// over several lanes; every row of the tile belongs to exactly one call.
// Fragment payloads: when a fragment becomes its pixel's current minimum,
// it records its shading inputs -- its progress u along the span and the
// row's side progresses pl, pr -- in three 64-bit slots of the pixel, each
// tagged with its draw sequence and the half and side-order flags (an
// aligned 64-bit LDS store is single-copy atomic).  The pixel's final
// winner became the minimum when it ran and wrote all three; a later store
// by a fragment it beat carries another tag.  So where all three tags name
// the final winner, the payload is the winner's own, and the shading needs
// no replay of its setup, rows or span; elsewhere (a race, a sequence of
// 2^30 or more) the pixel is shaded the long way.
// Tag: the sequence (25 bits), then the sub-triangle's vertex order in
// fill_triangle_2 as (index of S.v[0], of S.v[1]) among the triangle's
// three vertices, 3 * p0 + p1 (15: a clip vertex is among them), whether it
// is inverted, the half and the side order.
#define SW_SLOT_SEQ 0x01FFFFFFu
#define SW_SLOT_NOPERM 15u
__device__ __forceinline__ uint32_t slot_code(const Setup &S, uint32_t i0, uint32_t i1, uint32_t i2)
{
    if (S.v[0].clip | S.v[1].clip | S.v[2].clip)
        return SW_SLOT_NOPERM | (S.inverted ? 16u : 0u);
    const uint32_t p0 = S.v[0].a == i0 ? 0u : S.v[0].a == i1 ? 1u : 2u;
    const uint32_t p1 = S.v[1].a == i0 ? 0u : S.v[1].a == i1 ? 1u : 2u;
    if (p0 == p1)   // repeated vertex ids: no single order to record
        return SW_SLOT_NOPERM | (S.inverted ? 16u : 0u);
    return 3u * p0 + p1 + (S.inverted ? 16u : 0u);
}
__device__ __forceinline__ void put_slots(unsigned long long *slots, int n, int i, uint32_t seq, uint32_t code,
                                          bool lower, bool lr, float u, float pl, float pr)
{
    if (seq > SW_SLOT_SEQ)
        return;
    const uint64_t tag = (uint64_t)(seq | code << 25 | (lower ? 1u << 30 : 0u) | (lr ? 1u << 31 : 0u)) << 32;
    __hip_atomic_store(&slots[i], tag | __float_as_uint(u), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_store(&slots[n + i], tag | __float_as_uint(pl), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_store(&slots[2 * n + i], tag | __float_as_uint(pr), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_WORKGROUP);
}

// c ? a : b, field by field: a select of whole structs would take their
// addresses and put the setup in private memory.
__device__ __forceinline__ Side pick(bool c, const Side &a, const Side &b)
{
    Side r;
    r.ratio = c ? a.ratio : b.ratio;
    r.x = c ? a.x : b.x;
    r.it.ta = c ? a.it.ta : b.it.ta;
    r.it.ts = c ? a.it.ts : b.it.ts;
    r.it.ba = c ? a.it.ba : b.it.ba;
    r.it.bs = c ? a.it.bs : b.it.bs;
    r.it.v0 = c ? a.it.v0 : b.it.v0;
    r.it.v1 = c ? a.it.v1 : b.it.v1;
    return r;
}

// Calls row(y, R) for every row y in [wy0, wy1) that fill_triangle_2 draws
// for S (renderer.cpp:462-558), R holding the row's sides, half and side
// order.  The upper half's rows [up_y0, up_y1) and the lower half's
// [lo_y0, lo_y1) are adjacent when both exist (up_y1 == lo_y0 == iy1), so
// one loop walks them: the long side has taken y - ysL steps at row y in
// either half (row_state's two counts agree), a short side y minus its
// half's first row, and at the switch the lower short side starts from its
// setup state -- the same adds in the same order as a per-half replay,
// with every lane of a wave in one loop over its own rows.
template <typename RowF>
__device__ __forceinline__ void walk_rows(const Setup &S, int wy0, int wy1, RowF row)
{
    const bool up = S.has_up && S.up_y1 > S.up_y0, lo = S.has_lo && S.lo_y1 > S.lo_y0;
    if (!up && !lo)
        return;
    const int ya = max(wy0, up ? S.up_y0 : S.lo_y0), yb = min(wy1, lo ? S.lo_y1 : S.up_y1);
    if (ya >= yb)
        return;
    Row R;
    R.lower = !(up && ya < S.up_y1);
    {
        Side lng = S.L, sh = pick(R.lower, S.Sl, S.Su);
        lng.steps(ya - S.ysL);
        sh.steps(ya - (R.lower ? S.lo_y0 : S.up_y0));
        R.lr = R.lower ? S.lr_lo : S.lr_up;
        R.left = pick(R.lr, sh, lng);
        R.right = pick(R.lr, lng, sh);
    }
    for (int y = ya; y < yb; y++) {
        if (!R.lower && y == S.up_y1) {   // into the lower half
            const Side lng = pick(R.lr, R.right, R.left);
            R.lower = true;
            R.lr = S.lr_lo;
            R.left = pick(R.lr, S.Sl, lng);
            R.right = pick(R.lr, lng, S.Sl);
        }
        row(y, R);
        R.left.steps(1);
        R.right.steps(1);
    }
}

// check_mark: t came from a list that holds unmarked triangles too (the
// marks are read beside the vertices, the triangle skipped when unmarked).
__device__ void raster_depth(const JobDev &J, uint32_t t, TileRect tr, int wy0, int wy1,
                             unsigned long long *keys, int TW, uint32_t &frags, uint32_t &nans,
                             unsigned long long *slots = nullptr, int nslot = 0, bool check_mark = false)
{
    SubTri st[2];
    const uint32_t i0 = ro(J.tri)[3 * t], i1 = ro(J.tri)[3 * t + 1], i2 = ro(J.tri)[3 * t + 2];
    if (check_mark && !((ro(J.yes)[i0] != 0) & (ro(J.yes)[i1] != 0) & (ro(J.yes)[i2] != 0)))
        return;
    int ns = tri_subs_v<true>(J, i0, i1, i2, st);
#pragma unroll
    for (int s = 0; s < 2; s++) {
        Setup S;
        if (s >= ns || !setup_subtri(st[s], J.V, S))
            continue;
        const uint32_t seq = 2u * t + (uint32_t)s;
        const uint32_t code = slots ? slot_code(S, i0, i1, i2) : 0u;
        walk_rows(S, wy0, wy1, [&](int y, const Row &R) {
            Span sp;
            if (!span_of(R, J.V, sp))
                return;
            int xa = max(sp.x1, tr.x0), xb = min(sp.x2, tr.x1);
            if (xa >= xb)
                return;
            Interp q = sp.q;
            q.steps(xa - sp.x1);
            const int rowi = (y - tr.y0) * TW - tr.x0;
            unsigned long long *row = keys + rowi;
            for (int x = xa; x < xb; x++, q.step()) {
                const float u = q.u();
                float z = q.value(u);
                if (z < SW_NEAR)
                    continue;
                frags++;
                if (z != z) {
                    nans++;
                    continue;
                }
                if (__float_as_uint(z) >= SW_MAX_Z_BITS)
                    continue;
                const unsigned long long key = vis_key(z, seq);
                if (slots) {
                    if (key < atomicMin(&row[x], key))
                        put_slots(slots, nslot, rowi + x, seq, code, R.lower, R.lr, u, sp.pl, sp.pr);
                } else {
                    atomicMin(&row[x], key);
                }
            }
        });
    }
}

// Shading attributes of a sub-triangle vertex (v_world, normal_world, raw tex
// coords, the last in the fourth lanes of world and nworld); clip vertices
// are rebuilt as in renderer.cpp:290-297.
__device__ void vertex_attr(const JobDev &J, const NodeDev &N, const SVert &v, f3 &w, f3 &n, f2 &t)
{
    if (!v.clip) {
        const float4 w4 = ro4(J.world, v.a), n4 = ro4(J.nworld, v.a);
        w = f3{w4.x, w4.y, w4.z};
        n = f3{n4.x, n4.y, n4.z};
        t = f2{w4.w, n4.w};
        return;
    }
    const float4 wa4 = ro4(J.world, v.a), wb4 = ro4(J.world, v.b);
    const f3 wa = {wa4.x, wa4.y, wa4.z}, wb = {wb4.x, wb4.y, wb4.z};
    w = wa + (wb - wa) * v.cut;
    f2 ta = {wa4.w, ro4(J.nworld, v.a).w}, tb = {wb4.w, ro4(J.nworld, v.b).w};
    t = ta + (tb - ta) * v.cut;
    f3 na = {ro(J.nx)[v.a], ro(J.ny)[v.a], ro(J.nz)[v.a]}, nb = {ro(J.nx)[v.b], ro(J.ny)[v.b], ro(J.nz)[v.b]};
    f3 nobj = (na.x == nb.x && na.y == nb.y && na.z == nb.z) ? na : normalize(na + (nb - na) * v.cut);
    n = normal_world(N.RS, nobj);
}

// The job's view of primitive P's texture (pixel_shaders.cpp:288-302).
__device__ __forceinline__ TexView tex_view(const JobDev &J, const PrimDev &P)
{
    TexView X{nullptr, P.color, 1, 1};
    if (P.tex >= 0) {
        const TexDev T = ro(J.texs)[P.tex];
        X.p = J.texels + T.offset;
        X.w = T.w;
        X.h = T.h;
    }
    return X;
}

// GENERIC: the job's shader is any SWEGL_SHADER_* kind, decoded into a light
// part (0 none, 1 flat, 2 phong) and a texture part (0 the material colour,
// 1 nearest, 2 bilinear); otherwise it is SWEGL_SHADER_PHONG_BILINEAR, the
// kernel variant the common case runs without the other branches.
// Shading of triangle t's fragment with progress u on the row and span
// already set up (S, R, sp).
template <bool GENERIC>
__device__ uint32_t shade_from(const JobDev &J, uint32_t t, const Setup &S, const Row &R, const Span &sp,
                               float u, uint32_t &oob)
{
    int lk = 2, tk = 2;
    bool combined = true;
    if (GENERIC) {
        const int k = J.shader;
        lk = (k == 1 || k == 2 || k == 9) ? 2 : (k == 3 || k == 4 || k == 8) ? 1 : 0;
        tk = (k == 1 || k == 3 || k == 7) ? 2 : (k == 2 || k == 4 || k == 6) ? 1 : 0;
        combined = k >= 1 && k <= 4;
    }
    const PrimDev P = ro(J.prims)[prim_of_tri(J, t)];
    const NodeDev N = ro(J.nodes)[P.node];
    const TexView X = tex_view(J, P);
    const int tw = X.w, th = X.h;
    TriAttr A;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        f3 w, n;
        f2 tc;
        vertex_attr(J, N, S.v[k], w, n, tc);
        A.w[k] = w;
        A.n[k] = S.inverted ? neg(n) : n;
        A.t[k] = f2{tc.x * (float)tw, tc.y * (float)th};
    }
    ScanShade SS = prepare_scanline(A, R.lower, R.lr, sp.pl, sp.pr);
    if (!GENERIC) {
        uint32_t color = bilinear(X, SS.t_left + SS.t_dir * u, oob);
        f3 center = SS.v + SS.vdir * u;
        f3 normal = normalize(SS.n + SS.ndir * u);
        int32_t light = phong(center, normal, J.V.cam, J.ambient, J.sun, J.sun_I, J.lights, (int)J.n_lights);
        return combine(color, light);
    }
    uint32_t color = P.color;   // pixel_shader_t::color (default material for material_id -1)
    if (tk == 1)
        color = nearest(X, SS.t_left + SS.t_dir * u);
    else if (tk == 2)
        color = bilinear(X, SS.t_left + SS.t_dir * u, oob);
    if (lk == 0)
        return color;
    int32_t light;
    if (lk == 1) {
        light = flat_light(A.w[0], A.w[1], A.w[2], S.inverted, J.V.cam, J.ambient, J.sun, J.sun_I, J.lights,
                           (int)J.n_lights);
    } else {
        f3 center = SS.v + SS.vdir * u;
        f3 normal = normalize(SS.n + SS.ndir * u);
        light = phong(center, normal, J.V.cam, J.ambient, J.sun, J.sun_I, J.lights, (int)J.n_lights);
    }
    return combined ? combine(color, light) : (uint32_t)light;
}

// The colour of triangle t's fragment before lighting: the material colour,
// nearest texel or bilinear texel of the shader's texture part.  Its alpha
// is the fragment's alpha for every shader whose light part only scales
// b, g, r (a10, pixel_shaders.hpp:121-179) -- all but the light-only ones.
template <bool GENERIC>
__device__ uint32_t texture_from(const JobDev &J, uint32_t t, const Setup &S, const Row &R, const Span &sp,
                                 float u, uint32_t &oob)
{
    int tk = 2;
    if (GENERIC) {
        const int k = J.shader;
        tk = (k == 1 || k == 3 || k == 7) ? 2 : (k == 2 || k == 4 || k == 6) ? 1 : 0;
    }
    const PrimDev P = ro(J.prims)[prim_of_tri(J, t)];
    if (tk == 0)
        return P.color;
    const NodeDev N = ro(J.nodes)[P.node];
    const TexView X = tex_view(J, P);
    const int tw = X.w, th = X.h;
    TriAttr A;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        f3 w, n;
        f2 tc;
        vertex_attr(J, N, S.v[k], w, n, tc);
        A.w[k] = w;
        A.n[k] = n;
        A.t[k] = f2{tc.x * (float)tw, tc.y * (float)th};
    }
    ScanShade SS = prepare_scanline(A, R.lower, R.lr, sp.pl, sp.pr);
    if (tk == 1)
        return nearest(X, SS.t_left + SS.t_dir * u);
    return bilinear(X, SS.t_left + SS.t_dir * u, oob);
}

// Shading of the winning draw sequence seq at pixel (x, y): the sub-triangle,
// row and span re-derived by the same replay that produced the winner.
template <bool GENERIC>
__device__ uint32_t shade_pixel(const JobDev &J, uint32_t seq, int x, int y, uint32_t &oob)
{
    const uint32_t t = seq >> 1, s = seq & 1u;
    SubTri st[2];
    Setup S;
    Row R;
    Span sp;
    // The winner was produced by this very replay, so every step below
    // succeeds; the guards only keep a corrupted key from looping.
    if ((int)s >= tri_subs<true>(J, t, st))
        return 0u;
    SubTri one = st[0];   // a copy, not a runtime index: st stays in registers
    if (s)
        one = st[1];
    if (!setup_subtri(one, J.V, S, y) || !row_state(S, y, R) || !span_of(R, J.V, sp) || x < sp.x1 || x >= sp.x2)
        return 0u;
    Interp q = sp.q;
    q.steps(x - sp.x1);
    return shade_from<GENERIC>(J, t, S, R, sp, q.u(), oob);
}

// Shading of the winning sequence seq from its recorded payload (put_slots):
// the sub-triangle's vertices in fill_triangle_2's order, the half and side
// order, the side progresses and the span progress u.
template <bool GENERIC>
__device__ uint32_t shade_recorded(const JobDev &J, uint32_t seq, uint32_t tag, float u, float pl, float pr,
                                   uint32_t &oob)
{
    const uint32_t t = seq >> 1, s = seq & 1u, code = (tag >> 25) & 31u, perm = code & 15u;
    Setup S;
    if (perm < 9u) {
        // fill_triangle_2's vertex order, recorded: no z- or y-sort to redo
        const uint32_t id[3] = {ro(J.tri)[3 * t], ro(J.tri)[3 * t + 1], ro(J.tri)[3 * t + 2]};
        const uint32_t p0 = perm / 3u, p1 = perm % 3u, p2 = 3u - p0 - p1;
        const uint32_t ids[3] = {id[p0], id[p1], id[p2]};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            S.v[k].a = S.v[k].b = ids[k];
            S.v[k].cut = 0.0f;
            S.v[k].clip = 0;
        }
        S.inverted = (code & 16u) != 0;
    } else {
        SubTri st[2];
        if ((int)s >= tri_subs<true>(J, t, st))
            return 0u;   // not reached: the winner came out of this sub-triangle
        SubTri one = st[0];
        if (s)
            one = st[1];
        sort_subtri(one, S);
    }
    Row R;
    R.lower = (tag >> 30) & 1u;
    R.lr = (tag >> 31) & 1u;
    Span sp;
    sp.pl = pl;
    sp.pr = pr;
    return shade_from<GENERIC>(J, t, S, R, sp, u, oob);
}

template <int TW, int TH, bool GENERIC>
__device__ __forceinline__ void raster_tile(const BatchDev &B, const JobDev *__restrict__ jobs, uint32_t g,
                                            unsigned long long *keys, uint32_t *s_cnt, unsigned long long *slots)
{
    constexpr int NSLOT = TW * TH;
    const int jidx = find_job<JTILE>(B, jobs, g);
    const JobDev &J = jobs[jidx];
    const int lt = (int)(g - J.tile_base);
    const int tx = lt % J.tiles_x, ty = lt / J.tiles_x;
    TileRect tr;
    tr.x0 = J.V.vx + tx * TW;
    tr.y0 = J.V.vy + ty * TH;
    tr.x1 = min(tr.x0 + TW, J.V.vx + J.V.vw);
    tr.y1 = min(tr.y0 + TH, J.V.vy + J.V.vh);

    for (int i = threadIdx.x; i < TW * TH; i += 256)
        keys[i] = SW_KEY_EMPTY;
    __syncthreads();

    uint32_t frags = 0, nans = 0, oob = 0;
    // Replay the tile's binned triangles, then the job's big triangles whose
    // rectangle touches this tile.  With more than 256 of them, one lane per
    // triangle.  With fewer, a triangle's rows in the tile are split into
    // chunks of about 64 pixels (from its screen bounding box) and the chunks
    // are dealt to the lanes, so a large triangle does not leave one lane
    // looping over its pixels while the others idle.
    __shared__ uint32_t s_tri[256], s_rows[256], s_first[256], s_wsum[4];
    // The tile's entries: placed by k_front in its job's fixed layout
    // (every triangle whose rectangle holds the tile; marks are checked
    // here), or, when a tile of the job passed its capacity, by k_bin_fill
    // at the scanned offset (marked triangles only; the count also held
    // unmarked ones); then the job's big list.
    const bool fixed = B.fixed != 0 && B.ctr[SW_CTR_BIG0 + B.n_jobs + jidx] == 0;
    uint32_t nb;
    const uint32_t *bins;
    if (fixed) {
        nb = min(B.tile_count[g * SW_TILE_PAD], J.tile_cap);
        bins = B.bins_fixed + J.bin_base + (size_t)lt * J.tile_cap;
    } else {
        nb = B.tile_cursor[g * SW_TILE_PAD];
        bins = B.bins + B.tile_offset[g];
    }
    const uint32_t nbig = B.ctr[SW_CTR_BIG0 + jidx];
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    if (nb + nbig > 256) {
        // enough triangles to fill the lanes: one lane per triangle, all rows
        for (uint32_t e = threadIdx.x; e < nb + nbig; e += 256) {
            uint32_t t;
            bool check = false;
            if (e < nb) {
                t = ro(bins)[e];
                check = (t & SW_BIN_CHECK) != 0;   // never set in the filled layout
                t &= ~SW_BIN_CHECK;
            } else {
                t = ro(J.big)[e - nb];
                const uint2 r = ro2(J.rect, t);
                if (tx < (int)(r.x & 0xFFFF) || tx > (int)(r.y & 0xFFFF) || ty < (int)(r.x >> 16) ||
                    ty > (int)(r.y >> 16) || !marked(J, t))
                    continue;
            }
            raster_depth(J, t, tr, tr.y0, tr.y1, keys, TW, frags, nans, slots, NSLOT, check);
        }
    } else {
    for (uint32_t base = 0; base < nb + nbig; base += 256) {
        const uint32_t e = base + threadIdx.x;
        uint32_t t = 0xFFFFFFFFu, nch = 0;
        int ry0 = tr.y0, ry1 = tr.y1;
        bool check = false;
        if (e < nb + nbig) {
            if (e < nb) {
                t = ro(bins)[e];
                check = (t & SW_BIN_CHECK) != 0;
                t &= ~SW_BIN_CHECK;
            } else {
                t = ro(J.big)[e - nb];
                const uint2 r = ro2(J.rect, t);
                if (tx < (int)(r.x & 0xFFFF) || tx > (int)(r.y & 0xFFFF) || ty < (int)(r.x >> 16) ||
                    ty > (int)(r.y >> 16) || !marked(J, t))
                    t = 0xFFFFFFFFu;
            }
        }
        if (t != 0xFFFFFFFFu) {
            // rows [ceil(ymin), ceil(ymax)) are exactly the rows fill_triangle_2 visits for an
            // unclipped triangle; clipped or non-finite ones keep all rows of the tile
            const uint32_t i0 = ro(J.tri)[3 * t], i1 = ro(J.tri)[3 * t + 1], i2 = ro(J.tri)[3 * t + 2];
            if (check && !((ro(J.yes)[i0] != 0) & (ro(J.yes)[i1] != 0) & (ro(J.yes)[i2] != 0)))
                t = 0xFFFFFFFFu;   // unmarked: no chunks
        }
        if (t != 0xFFFFFFFFu) {
            const uint32_t i0 = ro(J.tri)[3 * t], i1 = ro(J.tri)[3 * t + 1], i2 = ro(J.tri)[3 * t + 2];
            const f3 a = frustum_to_screen(J.V, ld3(J.ndc, i0)), b = frustum_to_screen(J.V, ld3(J.ndc, i1)),
                     c = frustum_to_screen(J.V, ld3(J.ndc, i2));
            const float big = 1073741824.0f;
            int w = tr.x1 - tr.x0;
            if (fabsf(a.y) < big && fabsf(b.y) < big && fabsf(c.y) < big && fabsf(a.x) < big &&
                fabsf(b.x) < big && fabsf(c.x) < big && fminf(a.z, fminf(b.z, c.z)) >= SW_NEAR) {
                ry0 = max(tr.y0, cvtt(ceilf(fminf(a.y, fminf(b.y, c.y)))));
                ry1 = min(tr.y1, cvtt(ceilf(fmaxf(a.y, fmaxf(b.y, c.y)))));
                const int xa = max(tr.x0, (int)floorf(fminf(a.x, fminf(b.x, c.x))));
                const int xb = min(tr.x1, (int)ceilf(fmaxf(a.x, fmaxf(b.x, c.x))) + 1);
                w = max(xb - xa, 1);
            }
            const int rows = ry1 - ry0;
            nch = rows > 0 ? (uint32_t)min(max(rows * w / 64, 1), rows) : 0u;
        }
        if (nch == 0)
            ry0 = ry1 = tr.y0;
        // exclusive prefix of the chunk counts over the block
        const uint32_t incl = wave_incl_scan(nch);
        if (lane == 63)
            s_wsum[wave] = incl;
        s_tri[threadIdx.x] = t;
        s_rows[threadIdx.x] = (uint32_t)(ry0 - tr.y0) | ((uint32_t)(ry1 - tr.y0) << 8) | (nch << 16);
        __syncthreads();
        uint32_t before = 0;
        for (int k = 0; k < wave; k++)
            before += s_wsum[k];
        const uint32_t total = s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
        s_first[threadIdx.x] = before + incl - nch;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < total; i += 256) {
            // the entry whose chunks contain item i: last k with s_first[k] <= i and chunks
            int lo = 0, hi = 255;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_first[mid] <= i)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            // s_first is non-decreasing and i < total, so entry lo has a chunk holding i
            // (an entry without chunks shares its s_first with the next entry)
            const uint32_t pk = s_rows[lo], n = pk >> 16, ch = i - s_first[lo];
            const int r0 = tr.y0 + (int)(pk & 0xFF), rows = (int)((pk >> 8) & 0xFF) - (int)(pk & 0xFF);
            const int ya = r0 + (int)((uint32_t)rows * ch / n), yb = r0 + (int)((uint32_t)rows * (ch + 1) / n);
            raster_depth(J, s_tri[lo], tr, ya, yb, keys, TW, frags, nans, slots, NSLOT);
        }
        __syncthreads();   // the LDS tables are rewritten by the next 256 entries
    }
    }
    __syncthreads();

    const float maxz = __uint_as_float(SW_MAX_Z_BITS);
    for (int i = threadIdx.x; i < TW * TH; i += 256) {
        const int x = tr.x0 + (i % TW), y = tr.y0 + (i / TW);
        if (x >= tr.x1 || y >= tr.y1)
            continue;
        const unsigned long long k = keys[i];
        uint32_t color = 0;
        float z = maxz;
        uint32_t seq = 0xFFFFFFFFu;
        if (k != SW_KEY_EMPTY) {
            seq = (uint32_t)k;
            z = __uint_as_float((uint32_t)(k >> 32));
            bool done = false;
            if (slots) {
                const unsigned long long a = slots[i], b = slots[NSLOT + i], c = slots[2 * NSLOT + i];
                const uint32_t tag = (uint32_t)(a >> 32);
                if (tag == (uint32_t)(b >> 32) && tag == (uint32_t)(c >> 32) && (tag & SW_SLOT_SEQ) == seq) {
                    color = shade_recorded<GENERIC>(J, seq, tag, __uint_as_float((uint32_t)a),
                                                    __uint_as_float((uint32_t)b), __uint_as_float((uint32_t)c), oob);
                    done = true;
                }
            }
            if (!done)
                color = shade_pixel<GENERIC>(J, seq, x, y, oob);
        }
        J.out_color[(int64_t)(y - J.out_oy) * J.out_pitch + (x - J.out_ox)] = color;
        const int64_t zo = (int64_t)(y - J.V.vy) * J.V.vw + (x - J.V.vx);
        J.zbuf[zo] = z;
        if (J.tid)
            J.tid[zo] = seq;
    }
    // The depth-of-field pass reads this code instead of z.  A loop of its
    // own keeps its registers out of the shading loop's peak.
    if (J.dofc) {
        const FDiv dof_by(J.fdepth - 1.0f);
        const bool dof_fast = J.fdepth - 1.0f >= 1.0f && J.fdepth - 1.0f <= 0x1p20f;
        for (int i = threadIdx.x; i < TW * TH; i += 256) {
            const int x = tr.x0 + (i % TW), y = tr.y0 + (i / TW);
            if (x >= tr.x1 || y >= tr.y1)
                continue;
            const float z = __uint_as_float((uint32_t)(keys[i] >> 32));   // max_z where empty
            const float bf = blur_factor(z, J.fdist, J.fdepth, (float)J.rmax, dof_by, dof_fast);
            const int radius = min(max(cvtt(bf), 0), J.rmax);   // bf is in [0, rmax] for finite inputs
            J.dofc[(int64_t)(y - J.V.vy) * J.V.vw + (x - J.V.vx)] = (uint8_t)((bf != 0.0f ? 1 : 0) | (radius << 1));
        }
    }
    if (B.stats) {
        if (frags)
            atomicAdd(&s_cnt[0], frags);
        if (nans)
            atomicAdd(&s_cnt[1], nans);
        if (oob)
            atomicAdd(&s_cnt[2], oob);
    }
}

// Persistent: each workgroup takes tiles from the batch's work queue until
// the queue passes the last tile, so tiles of many jobs share the machine.
// At most 128 VGPRs, so four waves fit on each SIMD (512 registers a lane);
// the shading path needs about that many and one more halves the overlap.
template <int TW, int TH, bool GENERIC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_raster(BatchDev B,
                                                                                     const JobDev *__restrict__ jobs)
{
    __shared__ unsigned long long keys[TW * TH];
    // fragment payload slots (put_slots) for tiles of at most 32x32 pixels:
    // 24 KB, which keeps four workgroups on a CU
    constexpr bool SLOTS = TW * TH <= 1024;
    __shared__ unsigned long long slots[SLOTS ? 3 * TW * TH : 1];
    __shared__ uint32_t s_cnt[3];
    __shared__ uint32_t s_tile;
    if (threadIdx.x < 3)
        s_cnt[threadIdx.x] = 0;
    // Tiles come from SW_RASTER_QUEUES queues over the claim order
    // (B.tile_order, largest bin counts first), queue k owning the places
    // k, k + 8, k + 16, ... so that every queue runs from costly tiles to
    // cheap ones, each counting on its own cache line: same-address atomics
    // are serialised where they execute, so one shared queue would make every
    // claim wait for all others.  A workgroup drains its home queue (block
    // index mod 8), then the others in turn; a queue seen exhausted is never
    // asked again, so every workgroup stops after at most n_tiles + 8 claims
    // in total.
    const uint32_t T = B.n_tiles, home = blockIdx.x % SW_RASTER_QUEUES;
    uint32_t tried = 0;   // thread 0 only: queues found exhausted
    for (;;) {
        if (threadIdx.x == 0) {
            uint32_t tile = 0xFFFFFFFFu;
            while (tried < SW_RASTER_QUEUES) {
                const uint32_t k = (home + tried) % SW_RASTER_QUEUES;
                const uint32_t cnt = T > k ? (T - 1u - k) / SW_RASTER_QUEUES + 1u : 0u;   // places of queue k
                uint32_t *q = &B.ctr[SW_CTR_QUEUE0 + 64 * k];
                // counters only grow, so a stale read never skips a live queue
                if (__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < cnt) {
                    const uint32_t i = atomicAdd(q, 1u);
                    if (i < cnt) {
                        tile = ro(B.tile_order)[k + SW_RASTER_QUEUES * i];
                        break;
                    }
                }
                tried++;
            }
            s_tile = tile;
        }
        __syncthreads();
        const uint32_t g = s_tile;
        if (g == 0xFFFFFFFFu)
            break;
        raster_tile<TW, TH, GENERIC>(B, jobs, g, keys, s_cnt, SLOTS ? slots : nullptr);
        __syncthreads();   // tile done with keys and s_tile before the next claim
    }
    __syncthreads();
    if (B.stats && threadIdx.x == 0) {
        if (s_cnt[0])
            stat_add(B, SW_ST_FRAGMENTS, s_cnt[0]);
        if (s_cnt[1])
            stat_add(B, SW_ST_NAN, s_cnt[1]);
        if (s_cnt[2])
            stat_add(B, SW_ST_OOB, s_cnt[2]);
    }
}

// ------------------------------------------------------- k_raster_layers
__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    return v;
}

// One pixel's base and layers (renderer.cpp:486-550)
struct LayerPixel {
    float bz, lz[SW_MAX_LAYERS];
    uint32_t bc, bseq, lc[SW_MAX_LAYERS];
};

// The z test, shading and base or layer update of one fragment at a pixel
// (renderer.cpp:489-550); `shade` computes the fragment's colour.
template <typename ShadeF>
__device__ __forceinline__ void layer_update(LayerPixel &P, float z, uint32_t seq, int K, bool deferred,
                                             uint32_t &frags, uint32_t &nans, ShadeF shade)
{
    constexpr int KMAX = SW_MAX_LAYERS;
    const float maxz = __uint_as_float(SW_MAX_Z_BITS);
    if (z < SW_NEAR)
        return;
    frags++;
    if (z != z) {
        nans++;
        return;
    }
    if (z >= P.bz)
        return;
    const uint32_t color = shade();
    // deferred: the slots keep the fragment's sequence, shaded at the end
    const uint32_t keep = deferred ? seq : color;
    // the first layer that is empty or nearer than z (renderer.cpp:500-505)
    int at = K;
#pragma unroll
    for (int k = KMAX - 1; k >= 0; k--)
        if (k < K && (P.lz[k] == maxz || P.lz[k] < z))
            at = k;
    if ((color >> 24) == 255) {
        // opaque: the new base; layers behind it drop out (:506-522)
        P.bz = z;
        P.bc = keep;
        P.bseq = seq;
        float nz[KMAX];
        uint32_t nc[KMAX];
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            const int from = k + at;
            nz[k] = maxz;
            nc[k] = 0u;
#pragma unroll
            for (int m = 0; m < KMAX; m++)
                if (m == from && m < K) {
                    nz[k] = P.lz[m];
                    nc[k] = P.lc[m];
                }
        }
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            P.lz[k] = nz[k];
            P.lc[k] = nc[k];
        }
        return;
    }
    // transparent: insert at `at` (:523-550)
    bool all_used = false;
#pragma unroll
    for (int k = 0; k < KMAX; k++)
        if (k == K - 1)
            all_used = P.lz[k] != maxz;
    float cz = z;
    uint32_t cc = keep;
    if (all_used) {
        // shift toward layer 0: the new fragment lands at at-1, layer 0 falls out
#pragma unroll
        for (int k = KMAX - 1; k >= 0; k--)
            if (k < at) {
                const float tz = P.lz[k];
                const uint32_t tc = P.lc[k];
                P.lz[k] = cz;
                P.lc[k] = cc;
                cz = tz;
                cc = tc;
            }
    } else {
        // shift away from layer 0 up to the first empty layer
        bool done = false;
#pragma unroll
        for (int k = 0; k < KMAX; k++)
            if (!done && k >= at && k < K) {
                const float tz = P.lz[k];
                const uint32_t tc = P.lc[k];
                P.lz[k] = cz;
                P.lc[k] = cc;
                cz = tz;
                cc = tc;
                done = cz == maxz;
            }
    }
}

// A chunk entry's sub-triangle staged in LDS: the vertices after
// fill_triangle's z-sort when the triangle is drawn unclipped (the common
// case); clipped triangles are re-derived by the lanes that need them.
struct LayerSub {
    float sx[3], sy[3], sz[3];
    uint32_t a[3];
    uint32_t kind;   // 0: not drawn, 1: one unclipped sub-triangle (+2: front), 4: re-derive
};

// Viewports with transparency layers (renderer.cpp:486-550).  The layer
// updates depend on the order of the fragments at a pixel, so each 8x8 tile
// (one wave) applies its triangles in draw order: the tile's bin entries
// and the job's big triangles whose rectangle holds the tile are gathered in
// LDS and sorted by triangle index (bitonic sort; the index is the draw
// order).  A tile with more than SW_LAYER_LIST such entries walks all of the
// job's triangles instead, a chunk at a time through their tile rectangles,
// compacted in order.  Entries go in chunks of SW_LAYER_CHUNK (layer_chunk):
// each lane
// walks one entry's rows over the tile (the costly, independent part), then
// each lane takes one pixel through the chunk's fragments in order -- z test
// against the base, shade, then the base write (opaque, alpha 255) with the
// farther layers dropped, or the layer insertion (transparent).  The base and
// the K layers of each pixel live in registers.  At the end the lane writes
// the base colour, depth and sequence and the layers flattened into layer 0
// (viewport.cpp:80-84); k_tflatten/k_tcopy finish the flatten and
// post_shader_t's copy (viewport.cpp:53-67, post_shaders.hpp:22-47).
#define SW_LAYER_LIST 1024
#define SW_LAYER_CHUNK 64
__device__ __forceinline__ uint32_t blend_colors(uint32_t back, uint32_t front)
{
    // colors.cpp:39-57: alpha from the front, integer alpha union, channels
    // mixed in double and truncated to 8 bits
    const int a = (int)(front >> 24), ba = (int)(back >> 24);
    const uint32_t na = (uint32_t)(255 - ((255 - a) * (255 - ba) / 255));
    if (a == 255)
        return (front & 0xFFFFFFu) | (na << 24);
    if (a == 0)
        return (back & 0xFFFFFFu) | (na << 24);
    uint32_t out = na << 24;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double v = (double)((back >> (8 * c)) & 0xFF) * ((256 - a) / 256.0) +
                         (double)((front >> (8 * c)) & 0xFF) * (a / 256.0);
        out |= (uint32_t)(uint8_t)(int)v << (8 * c);   // v in [0, 255]
    }
    return out;
}

// The fragments of triangle t's unclipped sub-triangle over the tile's rows
// [ty0, ty0 + TH) and columns [tx0, tx0 + TW): the same row walk and span
// stepping as raster_depth (renderer.cpp:462-558), writing each covered
// pixel's depth to z[py * TW + px] and its bit to the returned mask.  With
// `opaque`, also the bit of every covered pixel whose texture-part colour
// (texture_from: material colour, nearest or bilinear texel) has alpha 255,
// from the same row and span state a re-derivation for that pixel gives.
template <bool GENERIC>
__device__ uint64_t layer_walk(const JobDev &J, uint32_t t, const SubTri &st, int tx0, int ty0, float *z,
                               uint64_t *opaque)
{
    constexpr int TW = SW_LAYER_TW, TH = SW_LAYER_TH;
    Setup S;
    if (!setup_subtri(st, J.V, S))
        return 0ull;
    // texture part of the colour (texture_from)
    int tk = 2;
    if (GENERIC) {
        const int k = J.shader;
        tk = (k == 1 || k == 3 || k == 7) ? 2 : (k == 2 || k == 4 || k == 6) ? 1 : 0;
    }
    const PrimDev P = ro(J.prims)[prim_of_tri(J, t)];
    TexView X{nullptr, P.color, 1, 1};
    TriAttr A;
    uint64_t opq = 0ull;
    const bool textured = opaque != nullptr && tk != 0;
    if (textured) {
        X = tex_view(J, P);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            A.w[k] = f3{0.0f, 0.0f, 0.0f};
            A.n[k] = f3{0.0f, 0.0f, 0.0f};
            A.t[k] = f2{ro4(J.world, S.v[k].a).w * (float)X.w, ro4(J.nworld, S.v[k].a).w * (float)X.h};
        }
    }
    const bool flat_opaque = (P.color >> 24) == 255;   // tk == 0: the material colour
    const int wy0 = ty0, wy1 = min(ty0 + TH, J.V.vy + J.V.vh), wx1 = min(tx0 + TW, J.V.vx + J.V.vw);
    uint64_t mask = 0ull;
    walk_rows(S, wy0, wy1, [&](int y, const Row &R) {
        Span sp;
        if (!span_of(R, J.V, sp))
            return;
        const int xa = max(sp.x1, tx0), xb = min(sp.x2, wx1);
        if (xa >= xb)
            return;
        Interp q = sp.q;
        q.steps(xa - sp.x1);
        ScanShade SS;
        if (textured)
            SS = prepare_scanline(A, R.lower, R.lr, sp.pl, sp.pr);
        for (int x = xa; x < xb; x++, q.step()) {
            const int p = (y - ty0) * TW + (x - tx0);
            const float u = q.u();
            z[p] = q.value(u);
            mask |= 1ull << p;
            bool op = flat_opaque;
            if (textured) {
                const f2 tc = SS.t_left + SS.t_dir * u;
                uint32_t no_count = 0;   // texel reads are counted where the survivors are shaded
                op = tk == 1 ? (nearest(X, tc) >> 24) == 255
                             : bilinear_channel(bilinear_taps(X, tc, no_count), 3) == 255;
            }
            opq |= (uint64_t)op << p;
        }
    });
    if (opaque)
        *opaque = opq;
    return mask;
}

// Apply the entries list[0, n) (n <= SW_LAYER_CHUNK, in draw order) to the
// tile.  Phase one, one lane per entry: the depth of every pixel the entry
// covers goes to LDS (s_z, 64 depths per entry, and a coverage mask).
// Phase two, one lane per pixel: the entries in order, each covered pixel's
// fragment applied to the pixel's base and layers; a fragment that passes the
// z test is shaded after setting its sub-triangle up again for its row.
// Near-clipped triangles (two sub-triangles or clip vertices) are set up by
// the pixel lanes themselves.
template <bool GENERIC>
__device__ __forceinline__ void layer_chunk(const JobDev &J, const uint32_t *list, uint32_t n, LayerSub *s_sub,
                                            float *s_z, uint64_t *s_mask, uint64_t *s_opq, int tx0, int ty0,
                                            int px, int py, bool mine, int K, bool deferred, LayerPixel &P,
                                            uint32_t &frags, uint32_t &nans, uint32_t &oob)
{
    constexpr int TW = SW_LAYER_TW;
    if (threadIdx.x < n) {
        const uint32_t t = list[threadIdx.x];
        SubTri st[2];
        const int ns = tri_subs<true>(J, t, st);
        LayerSub &L = s_sub[threadIdx.x];
        uint64_t m = 0ull;
        if (ns == 1 && !st[0].v[0].clip && !st[0].v[1].clip && !st[0].v[2].clip) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                L.sx[k] = st[0].v[k].s.x;
                L.sy[k] = st[0].v[k].s.y;
                L.sz[k] = st[0].v[k].s.z;
                L.a[k] = st[0].v[k].a;
            }
            L.kind = st[0].front ? 3u : 1u;
            m = layer_walk<GENERIC>(J, t, st[0], tx0, ty0, s_z + threadIdx.x * 64,
                                    deferred ? s_opq + threadIdx.x : nullptr);
        } else {
            L.kind = ns == 0 ? 0u : 4u;
            m = ns == 0 ? 0ull : ~0ull;   // near-clipped: every pixel lane sets it up itself
        }
        s_mask[threadIdx.x] = m;
    }
    __syncthreads();
    const int x = tx0 + px, y = ty0 + py, p = py * TW + px;
    for (uint32_t i = 0; i < n && mine; i++) {
        if (!((s_mask[i] >> p) & 1ull))
            continue;
        const LayerSub &L = s_sub[i];
        const uint32_t tt = list[i];
        if (L.kind & 1u) {
            layer_update(P, s_z[i * 64 + p], 2u * tt, K, deferred, frags, nans, [&] {
                // deferred: only the alpha matters here, found by the walk
                if (deferred)
                    return ((s_opq[i] >> p) & 1ull) ? 0xFF000000u : 0u;
                SubTri st;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    st.v[k].s = f3{L.sx[k], L.sy[k], L.sz[k]};
                    st.v[k].a = st.v[k].b = L.a[k];
                    st.v[k].cut = 0.0f;
                    st.v[k].clip = 0;
                }
                st.front = (L.kind & 2u) != 0;
                Setup S;
                Row R;
                Span sp;
                // the walk above produced this fragment, so these succeed
                if (!setup_subtri(st, J.V, S, y) || !row_state(S, y, R) || !span_of(R, J.V, sp))
                    return 0u;
                Interp q = sp.q;
                q.steps(x - sp.x1);
                return shade_from<GENERIC>(J, tt, S, R, sp, q.u(), oob);
            });
        } else if (L.kind == 4u) {
            SubTri st[2];
            const int ns = tri_subs<true>(J, tt, st);
#pragma unroll
            for (int sub = 0; sub < 2; sub++) {   // unrolled: st is indexed by constants
                if (sub >= ns)
                    break;
                Setup S;
                Row R;
                Span sp;
                if (!setup_subtri(st[sub], J.V, S, y) || !row_state(S, y, R) || !span_of(R, J.V, sp) ||
                    x < sp.x1 || x >= sp.x2)
                    continue;
                Interp q = sp.q;
                q.steps(x - sp.x1);
                const float u = q.u();
                layer_update(P, q.value(u), 2u * tt + (uint32_t)sub, K, deferred, frags, nans, [&] {
                    uint32_t no_count = 0;
                    return deferred ? texture_from<GENERIC>(J, tt, S, R, sp, u, no_count)
                                    : shade_from<GENERIC>(J, tt, S, R, sp, u, oob);
                });
            }
        }
    }
    __syncthreads();   // s_sub, s_z and s_mask are rewritten by the next chunk
}

template <bool GENERIC>
__global__ __launch_bounds__(SW_LAYER_THREADS) void k_raster_layers(BatchDev B, const JobDev *__restrict__ jobs)
{
    constexpr int TW = SW_LAYER_TW, TH = SW_LAYER_TH, KMAX = SW_MAX_LAYERS, NT = SW_LAYER_THREADS;
    static_assert(TW * TH == 64 && NT == 64, "the layer raster works on 64-pixel tiles, one wave each");
    constexpr uint32_t CH = SW_LAYER_CHUNK;
    __shared__ uint32_t s_list[SW_LAYER_LIST], s_wcount[NT / 64];
    __shared__ LayerSub s_sub[CH];
    __shared__ float s_z[CH * 64];
    __shared__ uint64_t s_mask[CH], s_opq[CH];
    const uint32_t g = blockIdx.x;
    const int jidx = find_job<JTILE>(B, jobs, g);
    const JobDev &J = jobs[jidx];
    const int lt = (int)(g - J.tile_base);
    const int tx = lt % J.tiles_x, ty = lt / J.tiles_x;
    const int tx0 = J.V.vx + tx * TW, ty0 = J.V.vy + ty * TH;
    // this lane's pixel
    const int px = (int)(threadIdx.x % TW), py = (int)(threadIdx.x / TW);
    const bool mine = tx0 + px < J.V.vx + J.V.vw && ty0 + py < J.V.vy + J.V.vh;
    const int K = J.layers;
    // every shader but the light-only ones takes its alpha from the texture
    // part: only the alpha is needed in order, the survivors are shaded last
    const bool deferred = !GENERIC || (J.shader != 8 && J.shader != 9);   // 8, 9: lights_flat, lights_phong alone
    const float maxz = __uint_as_float(SW_MAX_Z_BITS);
    LayerPixel P;
    P.bz = maxz;
    P.bc = 0u;
    P.bseq = 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        P.lz[k] = maxz;
        P.lc[k] = 0u;
    }
    uint32_t frags = 0, nans = 0, oob = 0;
    const int lane = __lane_id();
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    // a drawn triangle (all three vertices marked) whose rectangle holds this tile
    auto covers = [&](uint32_t t) {
        const uint2 r = ro2(J.rect, t);
        return r.x != RECT_EMPTY && tx >= (int)(r.x & 0xFFFF) && tx <= (int)(r.y & 0xFFFF) &&
               ty >= (int)(r.x >> 16) && ty <= (int)(r.y >> 16) && marked(J, t);
    };
    const uint32_t nb = B.tile_cursor[g * SW_TILE_PAD], off = B.tile_offset[g];
    const uint32_t nbig = B.ctr[SW_CTR_BIG0 + jidx];
    // the tile's triangles: bin entries, then the job's big ones whose
    // rectangle holds this tile, while they fit in LDS
    bool listed = nb <= SW_LAYER_LIST;
    uint32_t total = 0;
    if (listed) {
        for (uint32_t e = threadIdx.x; e < nb; e += NT)
            s_list[e] = ro(B.bins)[off + e];
        total = nb;
        const int wave = (int)(threadIdx.x >> 6);
        for (uint32_t base = 0; base < nbig && listed; base += NT) {
            const uint32_t e = base + threadIdx.x;
            const uint32_t t = e < nbig ? ro(J.big)[e] : 0u;
            const bool cov = e < nbig && covers(t);
            const uint64_t bal = __ballot(cov);
            if (lane == 0)
                s_wcount[wave] = (uint32_t)__popcll(bal);
            __syncthreads();
            uint32_t before = 0, sum = 0;
            for (int k = 0; k < NT / 64; k++) {
                before += k < wave ? s_wcount[k] : 0u;
                sum += s_wcount[k];
            }
            const uint32_t at = total + before + (uint32_t)__popcll(bal & lt_mask);
            if (cov && at < SW_LAYER_LIST)
                s_list[at] = t;
            total += sum;
            listed = total <= SW_LAYER_LIST;   // the same for every lane
            __syncthreads();   // s_wcount is rewritten by the next pass
        }
    }
    if (listed) {
        // bitonic sort of the first n2 >= total entries (padding sorts last)
        uint32_t n2 = 1;
        while (n2 < total)
            n2 <<= 1;
        for (uint32_t e = total + threadIdx.x; e < n2; e += NT)
            s_list[e] = 0xFFFFFFFFu;
        __syncthreads();
        for (uint32_t k = 2; k <= n2; k <<= 1)
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t i = threadIdx.x; i < n2; i += NT) {
                    const uint32_t o = i ^ j;
                    if (o > i) {
                        const uint32_t a = s_list[i], b = s_list[o];
                        if ((a > b) == ((i & k) == 0)) {
                            s_list[i] = b;
                            s_list[o] = a;
                        }
                    }
                }
                __syncthreads();
            }
        for (uint32_t c = 0; c < total; c += CH)
            layer_chunk<GENERIC>(J, s_list + c, min(CH, total - c), s_sub, s_z, s_mask, s_opq, tx0, ty0, px, py, mine, K,
                                 deferred, P, frags, nans, oob);
    } else {
        // too many for LDS: every triangle of the job in draw order, CH at a time
        const int wave = (int)(threadIdx.x >> 6);
        for (uint32_t base = 0; base < J.n_tris; base += CH) {
            const uint32_t t = base + threadIdx.x;
            const bool cov = threadIdx.x < CH && t < J.n_tris && covers(t);
            const uint64_t bal = __ballot(cov);
            if (lane == 0)
                s_wcount[wave] = (uint32_t)__popcll(bal);
            __syncthreads();
            uint32_t before = 0, total = 0;
            for (int k = 0; k < NT / 64; k++) {
                before += k < wave ? s_wcount[k] : 0u;
                total += s_wcount[k];
            }
            if (cov)
                s_list[before + (uint32_t)__popcll(bal & lt_mask)] = t;
            __syncthreads();
            layer_chunk<GENERIC>(J, s_list, total, s_sub, s_z, s_mask, s_opq, tx0, ty0, px, py, mine, K, deferred, P,
                                 frags, nans, oob);
        }
    }
    if (mine && deferred) {
        // shade the base and the surviving layers (sequences in the slots)
        const int x = tx0 + px, y = ty0 + py;
        P.bc = P.bseq != 0xFFFFFFFFu ? shade_pixel<GENERIC>(J, P.bseq, x, y, oob) : 0u;
#pragma unroll
        for (int k = 0; k < KMAX; k++)
            if (k < K)
                P.lc[k] = P.lz[k] != maxz ? shade_pixel<GENERIC>(J, P.lc[k], x, y, oob) : 0u;
    }
    if (mine) {
        const int x = tx0 + px, y = ty0 + py;
        J.out_color[(int64_t)(y - J.out_oy) * J.out_pitch + (x - J.out_ox)] = P.bc;
        const int64_t zo = (int64_t)(y - J.V.vy) * J.V.vw + (x - J.V.vx);
        J.zbuf[zo] = P.bz;
        if (J.tid)
            J.tid[zo] = P.bseq;
        // layers 1..K-1 flattened into layer 0 (viewport.cpp:43-51, 80-82)
        uint32_t f = P.lc[0];
#pragma unroll
        for (int k = 1; k < KMAX; k++)
            if (k < K && (P.lc[k] >> 24) != 0)
                f = blend_colors(f, P.lc[k]);
        J.tflat[zo] = f;
    }
    if (B.stats) {
        // a wave sum, one atomic per counter and wave
        const uint32_t f = wave_sum(frags), nn = wave_sum(nans), o = wave_sum(oob);
        if (lane == 0) {
            if (f)
                stat_add(B, SW_ST_FRAGMENTS, f);
            if (nn)
                stat_add(B, SW_ST_NAN, nn);
            if (o)
                stat_add(B, SW_ST_OOB, o);
        }
    }
}

// viewport_t::flatten(layer 0) (viewport.cpp:53-67): layer 0 over the screen
// pixel at (i, j) from the surface's origin -- the reference ignores the
// viewport's offset here -- or the screen pixel where layer 0 is empty.
__global__ __launch_bounds__(256) void k_tflatten(BatchDev B, uint32_t job, const JobDev *__restrict__ jobs)
{
    const JobDev &J = jobs[job];
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= (uint32_t)(J.V.vw * J.V.vh))
        return;
    const int i = (int)(p % (uint32_t)J.V.vw), j = (int)(p / (uint32_t)J.V.vw);
    const uint32_t back = J.surface[(int64_t)j * J.surface_pitch + i];
    const uint32_t f = J.tflat[p];
    J.tflat[p] = (f >> 24) != 0 ? blend_colors(back, f) : back;
}
// post_shader_t::shade with layers (post_shaders.hpp:22-47): layer 0 to the
// screen rows y + j, columns from 0 (the reference ignores the x offset).
__global__ __launch_bounds__(256) void k_tcopy(BatchDev B, uint32_t job, const JobDev *__restrict__ jobs)
{
    const JobDev &J = jobs[job];
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= (uint32_t)(J.V.vw * J.V.vh))
        return;
    const int i = (int)(p % (uint32_t)J.V.vw), j = (int)(p / (uint32_t)J.V.vw);
    J.surface[(int64_t)(j + J.V.vy) * J.surface_pitch + i] = J.tflat[p];
}

// ---------------------------------------------------------------- k_dof
// Depth of field (build spec of post_shader_depth_box, post_shaders.hpp:51-131
// with remap_clipped, lerp.hpp:24-43): 64x16 output pixels per block, colour
// and blur factor of the tile plus an R_max halo staged in LDS.
// Exact floor(x / c) for 1 <= c <= 1024 and x < 2^18 (a tap count and a
// sum of 8-bit channels) as a high multiply by m = ceil(2^32 / c): the
// error x * (m * c - 2^32) / 2^32 < 2^18 * 2^10 / 2^32 stays below 1/c.
struct DofRecip {
    uint32_t m[4 * SW_MAX_DOF_RADIUS * SW_MAX_DOF_RADIUS + 1];
};
constexpr DofRecip make_dof_recip()
{
    DofRecip t{};
    for (uint32_t c = 2; c <= 4u * SW_MAX_DOF_RADIUS * SW_MAX_DOF_RADIUS; c++)
        t.m[c] = 0xFFFFFFFFu / c + 1u;
    return t;
}
__constant__ DofRecip k_dof_recip = make_dof_recip();
__device__ __forceinline__ uint32_t div_count(uint32_t x, uint32_t c)
{
    return c == 1 ? x : __umulhi(x, k_dof_recip.m[c]);
}

// LDS bytes of a k_dof<RMAX> workgroup: two tap words per cell of the
// output tile plus an RMAX halo on every side (20 KB for RMAX 8: eight
// workgroups on a CU).
__host__ __device__ inline size_t dof_lds_bytes(int RMAX)
{
    return (size_t)(SW_DOF_TW + 2 * RMAX) * (SW_DOF_TH + 2 * RMAX) * 8;
}

// 256 threads as 64 columns x 4 rows.  Consecutive workgroups run on
// different XCDs (round robin over 8), so the job's tiles are dealt out in
// 8 contiguous runs, one per XCD: neighbouring tiles, whose halos overlap,
// then share that XCD's L2.  The raster stored, per pixel, whether its blur
// factor is nonzero (an "unfocused" tap) and its clamped radius.
// RMAX bounds the wave's halo (the loads below are unrolled for it): the
// launch picks 8 when no job of the launch has a larger radius.
// Each LDS cell holds a tap as two words, (b, r) and (g, 1) as 16-bit
// halves for an unfocused pixel, zeros for a focused pixel or a cell
// outside the viewport (never an unfocused tap, which is what the
// reference's clipped loops give): a tap is one 64-bit LDS read and two
// adds, its count coming with the colour.
template <int RMAX>
__global__ __launch_bounds__(256) void k_dof(BatchDev B, const JobDev *__restrict__ jobs)
{
    extern __shared__ uint2 s_dyn[];
    const JobDev &J = jobs[find_job<JD>(B, jobs, blockIdx.x)];
    const uint32_t n = dof_blocks(J.V.vw, J.V.vh), q = n / 8;
    uint32_t lb = blockIdx.x - J.dblock0;
    if (lb < 8 * q)
        lb = (lb % 8) * q + lb / 8;
    const int bx = (J.V.vw + SW_DOF_TW - 1) / SW_DOF_TW;
    const int ox = (int)(lb % bx) * SW_DOF_TW, oy = (int)(lb / bx) * SW_DOF_TH;   // viewport-relative
    const int R = J.rmax;
    const int SW = SW_DOF_TW + 2 * R, SH = SW_DOF_TH + 2 * R;
    // The LDS window has the pitch of the largest halo (P) and SH rows; its
    // cells are dealt to the threads in order, NC / 256 each (the window of
    // the largest halo is a multiple of 256 cells), so no lane idles on a
    // ragged row.
    constexpr int P = SW_DOF_TW + 2 * RMAX, NC = P * (SW_DOF_TH + 2 * RMAX), PERT = NC / 256;
    static_assert(NC % 256 == 0, "the LDS window is dealt out in whole rounds of 256 cells");
    uint2 *s_tap = s_dyn;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    // Every load of the halo is issued before the first is used, so the
    // workgroup waits for memory once rather than once per row.  A cell
    // outside the viewport (or past the window's SW columns) loads the
    // tile's first pixel instead (no branch per cell) and is zeroed after.
    // Offsets are 32-bit: a surface holds at most 2^30 words.
    const SW_RO uint32_t *colr = ro(J.out_color);
    const SW_RO uint8_t *code = ro(J.dofc);
    const int cbase = (J.V.vy - J.out_oy) * J.out_pitch + (J.V.vx - J.out_ox);   // colour of viewport (0, 0)
    const int safe_c = cbase + oy * J.out_pitch + ox, safe_k = oy * J.V.vw + ox;
    uint32_t cv[PERT], kv[PERT];
    bool okv[PERT];
#pragma unroll
    for (int k = 0; k < PERT; k++) {
        const int i = (int)threadIdx.x + 256 * k, r = i / P, c = i % P;
        const int ly = oy - R + r, lx = ox - R + c;
        okv[k] = r < SH && c < SW && ly >= 0 && ly < J.V.vh && lx >= 0 && lx < J.V.vw;
        cv[k] = colr[okv[k] ? cbase + ly * J.out_pitch + lx : safe_c];   // the raster's output: read-only here
        kv[k] = code[okv[k] ? ly * J.V.vw + lx : safe_k];
    }
    // the output pixels' own colours and codes (in the rows just loaded:
    // cache hits)
    constexpr int PER = SW_DOF_TH / 4;
    uint32_t src[PER], own[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int x = ox + tx, y = oy + ty + 4 * j;
        const bool in = x < J.V.vw && y < J.V.vh;
        src[j] = colr[in ? cbase + y * J.out_pitch + x : safe_c];
        own[j] = code[in ? y * J.V.vw + x : safe_k];
    }
#pragma unroll
    for (int k = 0; k < PERT; k++) {
        const uint32_t w = okv[k] && (kv[k] & 1u) ? ((cv[k] & 0x00FFFFFFu) | (1u << 24)) : 0u;
        s_tap[threadIdx.x + 256 * k] = make_uint2(w & 0x00FF00FFu, (w >> 8) & 0x00FF00FFu);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int py = ty + 4 * j, px = tx;
        const int x = ox + px, y = oy + py;
        if (x >= J.V.vw || y >= J.V.vh)
            continue;
        const int c = (py + R) * P + (px + R);
        const int radius = (int)(own[j] >> 1);
        uint32_t out = src[j];
        if (radius != 0) {
            // With radius <= 8 a sum has at most 256 taps of 255, under 2^16:
            // blue and red add as the halves of one word, green and the
            // count as the halves of another.
            constexpr bool PACKED = RMAX <= 8;
            uint32_t br = 0, gc = 0, b = 0, g = 0, r = 0, count = 0;
            auto tap = [&](int k) {   // LDS cell k
                const uint2 w = s_tap[k];
                if (PACKED) {
                    br += w.x;
                    gc += w.y;
                } else {
                    b += w.x & 0xFFFFu;
                    r += w.x >> 16;
                    g += w.y & 0xFFFFu;
                    count += w.y >> 16;
                }
            };
            if (J.post == 1) {   // box [x-r, x+r) x [y-r, y+r)
                for (int a = -radius; a < radius; a++)
                    for (int e = -radius; e < radius; e++)
                        tap(c + a * P + e);
            } else {             // 3x3 taps at {-r, 0, r}
                const int dr = radius * P;
#pragma unroll
                for (int a = -1; a <= 1; a++)
#pragma unroll
                    for (int e = -1; e <= 1; e++)
                        tap(c + a * dr + e * radius);
            }
            if (PACKED) {
                b = br & 0xFFFFu;
                r = br >> 16;
                g = gc & 0xFFFFu;
                count = gc >> 16;
            }
            if (count)
                out = (uint32_t)(uint8_t)div_count(b, count) | ((uint32_t)(uint8_t)div_count(g, count) << 8) |
                      ((uint32_t)(uint8_t)div_count(r, count) << 16) | (255u << 24);
        }
        J.surface[(int64_t)(y + J.V.vy) * J.surface_pitch + (x + J.V.vx)] = out;
    }
}

__global__ void k_fill32(uint32_t *p, uint64_t n, uint32_t w)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = w;
}

// ----------------------------------------------------------- launchers
#define LAUNCH_CHECK(...)                                                      \
    do {                                                                       \
        hipLaunchKernelGGL(__VA_ARGS__);                                       \
        hipError_t e_ = hipGetLastError();                                     \
        if (e_ != hipSuccess)                                                  \
            return e_;                                                         \
    } while (0)

hipError_t launch_world(const WorldArgs &W, hipStream_t s)
{
    if (W.n_vertices)
        LAUNCH_CHECK(k_world, dim3((W.n_vertices + 255) / 256), dim3(256), 0, s, W);
    return hipSuccess;
}
hipError_t launch_vertex(const BatchDev &B, uint32_t blocks, uint32_t w0, uint32_t n_words, hipStream_t s)
{
    // at least one block: the kernel also zeroes the counters
    LAUNCH_CHECK(k_vertex, dim3(blocks ? blocks : 1), dim3(256), 0, s, B, w0, n_words, B.jobs);
    return hipSuccess;
}
hipError_t launch_front(const BatchDev &B, uint32_t blocks, hipStream_t s)
{
    // at least one block: the last block to finish scans the tile counts
    // into the bin offsets
    LAUNCH_CHECK(k_front, dim3(blocks ? blocks : 1), dim3(256), 0, s, B, B.jobs);
    return hipSuccess;
}
hipError_t launch_bin_fill(const BatchDev &B, uint32_t blocks, hipStream_t s, hipEvent_t done)
{
    // one more block than the fill needs: the first one writes the claim order
    if (!B.n_tiles)
        return done ? hipEventRecord(done, s) : hipSuccess;
    if (done) {
        // the kernel's own completion signals `done`: no marker after it
        hipExtLaunchKernelGGL(k_bin_fill, dim3(blocks + 1), dim3(256), 0, s, nullptr, done, 0, B, B.jobs);
        return hipGetLastError();
    }
    LAUNCH_CHECK(k_bin_fill, dim3(blocks + 1), dim3(256), 0, s, B, B.jobs);
    return hipSuccess;
}
template <bool G>
static hipError_t launch_raster_t(const BatchDev &B, int tw, int th, dim3 g, hipStream_t s)
{
    dim3 b(256);
    if (tw == 8 && th == 8)
        LAUNCH_CHECK((k_raster<8, 8, G>), g, b, 0, s, B, B.jobs);
    else if (tw == 16 && th == 16)
        LAUNCH_CHECK((k_raster<16, 16, G>), g, b, 0, s, B, B.jobs);
    else if (tw == 32 && th == 16)
        LAUNCH_CHECK((k_raster<32, 16, G>), g, b, 0, s, B, B.jobs);
    else if (tw == 32 && th == 32)
        LAUNCH_CHECK((k_raster<32, 32, G>), g, b, 0, s, B, B.jobs);
    else if (tw == 64 && th == 32)
        LAUNCH_CHECK((k_raster<64, 32, G>), g, b, 0, s, B, B.jobs);
    else if (tw == 64 && th == 64)
        LAUNCH_CHECK((k_raster<64, 64, G>), g, b, 0, s, B, B.jobs);
    else
        return hipErrorInvalidValue;
    return hipSuccess;
}
hipError_t launch_raster(const BatchDev &B, int tw, int th, uint32_t max_groups, bool generic, hipStream_t s)
{
    if (!B.n_tiles)
        return hipSuccess;
    const dim3 g(B.n_tiles < max_groups ? B.n_tiles : max_groups);
    return generic ? launch_raster_t<true>(B, tw, th, g, s) : launch_raster_t<false>(B, tw, th, g, s);
}

hipError_t launch_layers(const BatchDev &B, bool generic, uint32_t pixels, hipStream_t s)
{
    if (!B.n_tiles)
        return hipSuccess;
    // one block per layer tile
    if (generic)
        LAUNCH_CHECK(k_raster_layers<true>, dim3(B.n_tiles), dim3(SW_LAYER_THREADS), 0, s, B, B.jobs);
    else
        LAUNCH_CHECK(k_raster_layers<false>, dim3(B.n_tiles), dim3(SW_LAYER_THREADS), 0, s, B, B.jobs);
    const uint32_t blocks = (pixels + 255) / 256;
    LAUNCH_CHECK(k_tflatten, dim3(blocks), dim3(256), 0, s, B, 0u, B.jobs);
    LAUNCH_CHECK(k_tcopy, dim3(blocks), dim3(256), 0, s, B, 0u, B.jobs);
    return hipSuccess;
}
hipError_t launch_dof(const BatchDev &B, uint32_t blocks, int max_radius, hipStream_t s)
{
    if (!blocks)
        return hipSuccess;
    if (max_radius <= 8)
        LAUNCH_CHECK(k_dof<8>, dim3(blocks), dim3(256), dof_lds_bytes(8), s, B, B.jobs);
    else
        LAUNCH_CHECK(k_dof<SW_MAX_DOF_RADIUS>, dim3(blocks), dim3(256), dof_lds_bytes(SW_MAX_DOF_RADIUS), s, B,
                     B.jobs);
    return hipSuccess;
}

hipError_t resolve_kernels()
{
    const void *k[] = {
        (const void *)k_world, (const void *)k_vertex, (const void *)k_front, (const void *)k_bin_fill,
        (const void *)k_raster<8, 8, false>, (const void *)k_raster<16, 16, false>,
        (const void *)k_raster<32, 16, false>, (const void *)k_raster<32, 32, false>,
        (const void *)k_raster<64, 32, false>, (const void *)k_raster<64, 64, false>,
        (const void *)k_raster<8, 8, true>, (const void *)k_raster<16, 16, true>,
        (const void *)k_raster<32, 16, true>, (const void *)k_raster<32, 32, true>,
        (const void *)k_raster<64, 32, true>, (const void *)k_raster<64, 64, true>,
        (const void *)k_raster_layers<false>, (const void *)k_raster_layers<true>,
        (const void *)k_tflatten, (const void *)k_tcopy,
        (const void *)k_dof<8>, (const void *)k_dof<SW_MAX_DOF_RADIUS>, (const void *)k_fill32,
    };
    for (const void *f : k) {
        hipFuncAttributes a;
        hipError_t e = hipFuncGetAttributes(&a, f);
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

hipError_t launch_fill32(uint32_t *p, uint64_t n, uint32_t word, hipStream_t s)
{
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536)
        blocks = 65536;
    if (blocks)
        LAUNCH_CHECK(k_fill32, dim3((uint32_t)blocks), dim3(256), 0, s, p, n, word);
    return hipSuccess;
}

}  // namespace sw
