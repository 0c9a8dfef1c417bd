// jobs.h -- device-side descriptors shared by the host orchestration
// (capi.cpp) and the kernels (kernels.hip).
//
// HBM layout of a resident scene (SoA, all primitives concatenated in node
// order, then primitive order):
//   px,py,pz / nx,ny,nz / tu,tv : float[n_vertices]   mesh_vertex_t::v, normal, tex_coords
//   tri                          : uint32[3*n_tris]    vertex ids per triangle in draw
//                                                      order (strips/fans expanded with
//                                                      the reference's winding)
//   prims / nodes / texs         : small tables
//   texels                       : uint32 bitmaps of all textures, concatenated
//   world, nworld                : float4[n_vertices]  v_world / normal_world for the
//                                                      current node poses; .w holds the
//                                                      texture coordinate u / v
// Per viewport job:
//   ndc                          : float4[n_vertices]  frustum coordinates
//   yes                          : uint8[n_vertices]   visibility marks
//   rect                         : uint2[n_tris]       tile rectangle of each triangle
//   big                          : uint32[n_tris]      triangles covering more than
//                                                      BatchDev::big_tiles tiles
#pragma once
#include <stdint.h>
#include "core.h"
#include "shade.h"

namespace sw {

struct PrimDev {
    uint32_t first_vertex, n_vertices;
    uint32_t first_tri, n_tris;
    uint32_t node;
    int32_t tex;           // texture index, or -1: 1x1 bitmap of `color`
    uint32_t color;        // material colour (default bitmap, pixel_shaders.cpp:290-294)
    int32_t double_sided;
};

struct NodeDev {
    m34 M;     // original_to_world_matrix (rows 0..2)
    m34 RS;    // scale(rotation, scale) for normals (vertex_shaders.hpp:63)
};

struct TexDev {
    uint64_t offset;   // into texels
    int32_t w, h;
};

#define SW_MAX_DOF_RADIUS 16
#define SW_MAX_LAYERS 4   // transparency layers per viewport (SWEGL_MAX_TRANSPARENCY_LAYERS)
// tile of the in-order layer raster: 64 pixels, one wave
#define SW_LAYER_TW 8
#define SW_LAYER_TH 8
#define SW_LAYER_THREADS 64
// Bin limit of layered waves: on 8x8 tiles a smaller limit would send most
// triangles of a close-up mesh to the big list, which every tile scans.
#define SW_LAYER_BIG_TILES 32
// Depth-of-field output tile (pixels) of one 256-thread workgroup.
#define SW_DOF_TW 64
#define SW_DOF_TH 16
__host__ __device__ inline uint32_t dof_blocks(int w, int h)
{
    return (uint32_t)(((w + SW_DOF_TW - 1) / SW_DOF_TW) * ((h + SW_DOF_TH - 1) / SW_DOF_TH));
}

struct JobDev {
    // ---- resident scene
    const float *px, *py, *pz, *nx, *ny, *nz, *tu, *tv;
    const uint32_t *tri;
    const PrimDev *prims;
    const NodeDev *nodes;
    const uint32_t *texels;
    const TexDev *texs;
    const Light4 *lights;
    uint32_t n_prims, n_lights, n_vertices, n_tris;
    float ambient, sun_I;
    f3 sun;
    const float4 *world, *nworld;
    // ---- per viewport
    float4 *ndc;
    uint8_t *yes;
    uint2 *rect;
    uint32_t *big;
    ViewXf V;
    uint32_t *out_color;      // raster colour target
    int32_t out_pitch;        // words per row
    int32_t out_ox, out_oy;   // absolute pixel (x,y) maps to out_color[(y-oy)*pitch + (x-ox)]
    uint32_t *surface;        // final surface (SDL_Surface pixels)
    int32_t surface_pitch;    // words per row
    float *zbuf;              // m_zbuffer, w*h, viewport relative
    uint32_t *tid;            // optional winning draw sequence, w*h
    int32_t shader;           // SWEGL_SHADER_*
    int32_t post;             // SWEGL_POST_*
    uint8_t *dofc;            // DoF: per pixel (blur factor != 0) | radius << 1, w*h
    int32_t layers;           // transparency layers (0: none)
    uint32_t *tflat;          // with layers: flattened layer colour per pixel, w*h
    float fdist, fdepth;
    int32_t rmax;
    // ---- tiles
    int32_t tw, th, tiles_x, tiles_y;
    uint32_t tile_base;       // first global tile id of this job
    // k_front's fixed layout (BatchDev::fixed): this job's tile k keeps its
    // first tile_cap entries at bins_fixed[bin_base + k * tile_cap + rank]
    uint32_t tile_cap;
    uint64_t bin_base;
    // ---- first block of this job in each batched launch
    uint32_t vblock0, tblock0, dblock0;
};

struct WorldArgs {
    const float *px, *py, *pz, *nx, *ny, *nz, *tu, *tv;
    const PrimDev *prims;
    const NodeDev *nodes;
    uint32_t n_prims, n_vertices;
    float4 *world, *nworld;
};

// Tile counters: tile g's bin count at tile_count[g * SW_TILE_PAD] and its
// fill rank (entries placed so far) one word later, a 128-byte line per tile, so the binning
// atomics of different tiles spread over all memory channels instead of
// queueing on the few that a dense array maps to.
#define SW_TILE_PAD 32

struct BatchDev {
    const JobDev *jobs;
    uint32_t n_jobs;
    uint32_t n_tiles;         // total tiles over all jobs
    uint32_t *tile_count;     // [n_tiles * SW_TILE_PAD], see SW_TILE_PAD
    uint32_t *tile_offset;    // [n_tiles + 1]
    uint32_t *tile_order;     // [n_tiles]: tiles by descending bin count (k_bin_fill's last block), the raster's claim order
    uint32_t *tile_cursor;    // tile_count + 1: the fill ranks
    uint32_t *bins;           // [capacity]: the counted, scanned and filled layout
    uint32_t *bins_fixed;     // the fixed layout's entries (JobDev::bin_base, tile_cap)
    uint32_t bin_capacity;
    uint32_t big_tiles;       // per-triangle bin limit; larger ones go to the big list
    uint32_t *ctr;            // counter words, layout below
    uint32_t stats;           // 1: accumulate the statistics counters
    // 1: k_front places every entry in its job's fixed layout as it counts
    // it (marks unchecked); only a job with a tile holding more entries than
    // its tile_cap (its word at SW_CTR_BIG0 + n_jobs + job set, and
    // SW_CTR_OVERFLOW) is filled by k_bin_fill into the counted layout,
    // after the offsets scan; 0: the counted layout for every job
    uint32_t fixed;
};

// Counter block: [0, 1024) statistics as 64 slots x 8 u64 (slot = block
// index mod 64, so no address takes one atomic per block; the host sums the
// slots), then the raster's tile queues (one per 256-byte line), the total
// bin entry count, the big-list length of each job j at SW_CTR_BIG0 + j,
// and its fixed-layout overflow word at SW_CTR_BIG0 + n_jobs + j.
#define SW_STAT_SLOTS 64
#define SW_CTR_STATS 0
#define SW_CTR_CONTROL 1024
#define SW_RASTER_QUEUES 8
#define SW_CTR_QUEUE0 SW_CTR_CONTROL                              // queue k at + 64k (own 256 B)
#define SW_CTR_TOTAL (SW_CTR_CONTROL + 64 * SW_RASTER_QUEUES)
// k_front's finished blocks: SW_TICKET_GROUPS counters of the blocks with
// index k mod SW_TICKET_GROUPS, then one of the finished groups, each on
// its own 128-byte line (same-address atomics serialise where they execute)
#define SW_TICKET_GROUPS 32
#define SW_CTR_OVERFLOW (SW_CTR_TOTAL + 32)   // nonzero: a tile passed tile_cap
// A fixed-layout bin entry with this bit set names a triangle that failed its
// own visibility test: it is drawn only if other triangles marked all three
// of its vertices, which the raster looks up (a job holds < 2^31 triangles)
#define SW_BIN_CHECK 0x80000000u
#define SW_CTR_TICKETS (SW_CTR_TOTAL + 64)
#define SW_CTR_BIG0 (SW_CTR_TICKETS + 32 * (SW_TICKET_GROUPS + 1))
// statistic indices within a slot
#define SW_ST_FRAGMENTS 0
#define SW_ST_NAN 1
#define SW_ST_DRAWN 2
#define SW_ST_BINS 3
#define SW_ST_OOB 4
#define SW_ST_BIG 5           // triangles routed to the big list

}  // namespace sw
