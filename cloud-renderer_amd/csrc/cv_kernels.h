// Shared declarations between the host runtime (cv_runtime.hip) and the
// gfx950 kernels (cv_kernels.hip).  Plain structs passed by value as kernel
// arguments; no pointers to host memory.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cv_mlp.h"

namespace cv {

constexpr int kMaxLevels = 16;
constexpr int kMaxSteps = 256;
constexpr int kTile = 8;           // default camera tile: 8x8 pixels = one wavefront
constexpr int kLdsSortMax = 16384; // billboards sorted in one workgroup's LDS
constexpr int kBatchSortMax = 8192; // billboards a batched LDS sort takes (64 KiB of keys)
constexpr int kFillChunk = 4096;    // elements per workgroup of the segmented fill
constexpr int kChordLdsWords = 4096; // largest LDS bitmap (bricks) per pass-1 tile for chord-fill aggregation
constexpr int kLightTile = 32;       // pass-1 work item: 32 x 32 light pixels of one billboard
constexpr unsigned kConeWaves = 6144; // persistent cone-trace wavefronts (256 CUs x 24)

// One mip level in the bricked layout: 4x4x4 bricks, texels Morton-ordered
// inside a brick (bit order x0 y0 z0 x1 y1 z1), bricks x-fastest.
struct LevelDesc {
    uint32_t n;      // texels per side
    uint32_t nb;     // bricks per side = ceil(n / 4)
    uint64_t base;   // first texel of the level (in texels)
    float fn;        // (float)n (the samplers' scale, without a per-sample conversion)
};

// Voxel grid of the volume (world bounds, index map)
struct GridDesc {
    float lo[3];     // world bounds low corner = position + bounds_lo
    float range[3];  // world bounds extent
    float inv_range[3];  // 1 / range, used only when range is a power of two
    float vpos[3];   // CloudVolume::position (billboard centres are relative)
    float stepSize;  // min(voxelSize)
    float invN;      // 1 / N, used only when N is a power of two
    int N;           // voxels per side
    uint32_t nb0;    // level-0 bricks per side
    int pow2_range;  // every range[a] is a power of two: x / range[a] == x * inv_range[a] exactly
    int pow2_N;      // N is a power of two: x / N == x * invN exactly
    int pow2_step;   // stepSize is a power of two: the chord counter i += stepSize equals k * stepSize
    float inv_step;  // 1 / stepSize
};

// Light view of pass 1 (Sun::update output + derived billboard basis)
struct LightView {
    float V[16];
    float S[3], U[3], B[3];
    float np[3];
    float clip;
    float lsx, lox, lsy, loy;  // x_view = xn * lsx + lox
    float znear, zfar;
    int W, H;
};

struct VolumeDesc {
    LevelDesc lv[kMaxLevels];
    int L;
    int format;    // CR_FORMAT_*
    int mode;      // CR_MODE_*
    int noise_octaves;
    uint32_t noise_seed;
    float noise_base_freq, noise_lacunarity, noise_gain;
    // the density noise of every level-0 texel (level-0 slot order; binary16
    // bits for RGBA16F volumes, float for RGBA32F), or null: computed in the build
    const void* ntex;
};

struct CameraView {
    float V[16];
    float S[3], U[3], B[3];
    float eye[3];
    float kx, ky;      // tan(fovy/2) * aspect, tan(fovy/2)
    float znear, zfar;
    int W, H;
    int tiles_x, tiles_y;
    int tile_w, tile_h;  // camera tile shape, tile_w * tile_h = 64 pixels = one wavefront
    int tile_shift;      // log2(tile_w)
};

struct ConeDesc {
    float h[kMaxSteps];    // cone height before step i (voxels)
    // textureLod's level and fraction at step i (lambda = lod + lodOffset,
    // GL 4.4 8.14.3): level llev[i], blended with llev[i] + 1 by lfrac[i]
    float lfrac[kMaxSteps];
    uint32_t llev4[kMaxSteps / 4];  // llev[i] = byte i % 4 of llev4[i / 4] (read with word loads)
    int steps;
    float tanHalf;
    float den;             // steps * downScaling
    float alpha_stop;
    float kscale;          // sigma * log2(e) / N
    int cone_mode;
    int count;             // march: count samples and two-level samples (counter runs only)
    int composite;
    int quantize;
    int compensated;
    float clear[4];
    float lpos[3];
    float ldir[3];         // normalize(direction) / N
    int directional;
    int pixfmt;
    uint64_t pitch;
    // learned far field (LEARNED kernels and segment generation)
    int learned_k;         // samples marched before the far field is predicted
    float err_bound;       // fall back when (1 - a_k) * predicted error > bound; < 0: always
    float feat_h[4];       // feature sampling heights (voxels) for LODs 3..6
    int feat_lvl[4];       // their levels, clamped to L - 1
    float sun[3];          // sun direction feature
    // soft-particle noise march (NM kernels): the ConeTraceShader noise
    // parameters, the octaveOffsets uniform and normalize(V's third row)
    int nm_octaves, nm_min_steps, nm_max_steps;
    float nm_step, nm_opacity, nm_freq_step, nm_pers_step, nm_adjust, nm_min_color, nm_color_scale;
    float nm_off[3];
    float nm_ray[3];
    // sun sprite (GL blend composite): view-space centre, colours, radii
    int sun_on;
    float sun_v[3];
    float sun_in[3], sun_out[3];
    float sun_r0, sun_r1;
};

// pass-1 work list: per-billboard pixel box and tile count, exclusive scan
// (offsets[S] = total), and (billboard, tile) items
// One pass-1 work item: a kLightTile^2 tile of a billboard's pixel box, the
// billboard's light-view data and the brick box of the chord slab behind the
// tile, prepared once per item by the work-list kernel.
struct RasterItem {
    int j;                    // billboard
    int x0, y0, x1, y1;       // pixel rectangle, inclusive
    int bx0, by0, bz0;        // slab brick box origin
    int nbx, nby, nbz;        // and extents
    int use_lds;              // the slab's words fit the LDS bitmap
    float cx, cy, cz, s;      // sphere centre (world) and radius
    float vx, vy;             // centre in light view
    float inv_tw;             // 1 / (x1 - x0 + 1)
    float inv_x, inv_xy;      // 1 / nbx, 1 / (nbx * nby) (bitmap flush decode)
    int slice, nslices;       // depth slice of the chords behind the tile (light_slices)
    int pad;
};
static_assert(sizeof(RasterItem) == 96, "work items are 96 bytes");

struct LightScratch {
    int4* lbox;
    int* counts;
    int* offsets;   // S + 1
    RasterItem* items;
    size_t items_cap;
    int words;      // LDS bitmap words of the raster kernel
};

// ---- launch wrappers (cv_kernels.hip) ----
void launch_fill_u64(uint64_t* p, uint64_t n, uint64_t value, hipStream_t s);
struct FillSeg {
    uint64_t* p;
    uint64_t n;
    uint64_t value;
};
void launch_fill_multi(const FillSeg* segs, int nseg, hipStream_t s);  // up to 4 segments, one launch
size_t light_items_bound(int S, float max_radius, const LightView& lv, const GridDesc& g, int chord);

// LDS bitmap words of the raster kernel for this grid (its WORDS)
__host__ __device__ inline int raster_words(const GridDesc& g) {
    const bool p2 = g.pow2_range && g.pow2_N && g.pow2_step;
    return p2 && g.nb0 <= 64 ? 2048 : 4096;
}
// The chord walk of a pass-1 item runs from the queue (the flat walk) when the
// grid has power-of-two extents and the chord fill is on.
__host__ __device__ inline bool raster_flat(const GridDesc& g, int chord) {
    return chord && g.pow2_range && g.pow2_N && g.pow2_step;
}
// Depth slices of a billboard's flat-walk items: the smallest power of two
// K <= 32 for which the brick box of a full tile's chord slab cut into K
// equal depth ranges fits the raster's LDS bitmap (a bound of the box
// slab_box computes: extent + 2 voxels of margin, +2 bricks of alignment);
// 1 when the chords are long enough that a slice's first sample index
// would not fit 16 bits.  Slices of one tile are separate work items.
__host__ __device__ inline int light_slices_box(const GridDesc& g, const LightView& lv, float s, int words) {
    if (!(s > 0.0f)) return 1;
    const float wu = fminf((float)kLightTile * (2.0f * lv.lsx / (float)lv.W), 2.0f * s);
    const float wv = fminf((float)kLightTile * (2.0f * lv.lsy / (float)lv.H), 2.0f * s);
    for (int K = 1; K < 32; K <<= 1) {
        const float dt = 2.0f * s / (float)K;
        long long w = 1;
        for (int a = 0; a < 3; ++a) {
            const float ext = (wu * fabsf(lv.S[a]) + wv * fabsf(lv.U[a])) + dt * fabsf(lv.B[a]);
            const float vox = g.range[a] / (float)g.N;
            long long nb = (long long)((ext / vox + 2.0f) * 0.25f) + 2;
            if (nb > (long long)g.nb0) nb = g.nb0;
            w *= nb;
        }
        if (w <= (long long)words) return K;
    }
    return 32;
}
__host__ __device__ inline int light_slices(const GridDesc& g, const LightView& lv, float s, int words) {
    if (2.0f * s / g.stepSize >= 65535.0f) return 1;
    return light_slices_box(g, lv, s, words);  // non-decreasing in s: the work-list bound takes the largest s
}
// the pass-1 work list (boxes, scan, items); launch_light_raster reads it
void launch_light_items(const GridDesc& g, const LightView& lv, int mode, const float4* boards, int S,
                        const LightScratch& ls, hipStream_t s);
void launch_light_raster(const GridDesc& g, const LightView& lv, int mode, const float4* boards, int S,
                         const LightScratch& ls, unsigned long long* keys, unsigned long long* fillbits,
                         unsigned long long* counters, hipStream_t s);
void launch_light_splat(const GridDesc& g, const LightView& lv, const float4* boards,
                        const unsigned long long* keys, float4* posmap, unsigned long long* shellbits,
                        unsigned long long* counters, hipStream_t s);
// zflags (optional, RGBA formats with N % 16 == 0): one byte per 16^3
// region of the level set, 1 when everything the region's build writes holds
// zero; regions that stay empty then write nothing
size_t build_zero_flags(const GridDesc& g, const VolumeDesc& vd);
// bytes of the density noise texture of an RGBA volume with noise (0: none)
size_t noise_tex_bytes(const GridDesc& g, const VolumeDesc& vd);
void launch_noise_tex(const GridDesc& g, const VolumeDesc& vd, void* ntex, hipStream_t s);
void launch_build_levels(const GridDesc& g, const VolumeDesc& vd, const unsigned long long* shellbits,
                         const unsigned long long* fillbits, void* levels, uint8_t* zflags, hipStream_t s);
void launch_mip(const VolumeDesc& vd, int k, void* levels, hipStream_t s);
void launch_cam_setup(const GridDesc& g, const CameraView& cv, const float4* boards, int S, float4* sph_world,
                      float4* sph_view, int4* sph_tiles, unsigned long long* keys, hipStream_t s);
size_t sort_scratch_words(int S);
// keys of a sort far to near from `point` alone (CloudVolume::sortBoards), and
// the identity order (billboards drawn in upload order)
void launch_sort_keys(const GridDesc& g, const float4* boards, int S, const float point[3], unsigned long long* keys,
                      hipStream_t s);
void launch_iota(int32_t* order, int S, hipStream_t s);
int launch_sort(unsigned long long* keys, int S, int32_t* order, int* tie_flag, hipStream_t s);
// camera tiles in the order the cone waves take them (heaviest first)
struct TileQueue {
    unsigned int* counter;
    const int* order;
};
size_t tile_scratch_ints(int ntiles);
TileQueue launch_bin(const int4* sph_tiles, const int32_t* order, int S, int tiles_x, int tiles_y, int words,
                     unsigned long long* masks, int* scratch, hipStream_t s);
// One scene of a batched cone-trace launch (device resident; the batched
// kernel reads it with wave-uniform loads).
struct BatchScene {
    GridDesc g;
    VolumeDesc vd;
    CameraView cv;
    ConeDesc cd;
    const void* levels;
    const float4* sph_world;
    const float4* sph_view;
    const int32_t* order;
    const unsigned long long* masks;
    void* image;
    int words;
    int tile0;  // first global tile of the scene
};

// One scene of a batch's voxelization and camera pass, as the batched stage
// kernels read it (device array, one entry per scene)
struct BatchVox {
    GridDesc g;
    LightView lv;
    VolumeDesc vd;
    CameraView cv;
    const float4* boards;
    int S;
    int chord;   // chord fill (rgba mode)
    int words;   // LDS bitmap words of the pass-1 raster
    int4* lbox;
    int* counts;
    int* offsets;
    RasterItem* items;
    int items_cap;
    unsigned long long* keys;  // pass-1 depth keys
    unsigned long long* shell;
    unsigned long long* fill;
    void* levels;
    uint8_t* zflags;
    int nlev;       // levels of the fused build
    uint32_t rpa;   // fused build regions per axis
    int tail_k;     // first level of the one-workgroup tail build (L: none)
    float4* sph_world;
    float4* sph_view;
    int4* sph_tiles;
    unsigned long long* sortkeys;
    int32_t* order;
    int* tie;
    int order_mode;  // CR_ORDER_*
    unsigned long long* masks;
    int mask_words;  // mask words per camera tile
    int* tile_count;
};
// A batched stage launch: workgroup b belongs to group member m with
// first[m] <= b < first[m + 1] (first[n] = total), which is scene scene[m].
struct BatchStage {
    const int* first;
    const int* scene;
    int n;
};
// one launch per stage over every scene of the batch that the batched path
// takes (format shared: FMT of the fused build); see batch_vox_eligible
enum BatchStageId { kBsBoxes, kBsScan, kBsExpand, kBsPrep, kBsRaster, kBsSplat, kBsBuild, kBsTail, kBsCam,
                    kBsSortIota, kBsSort, kBsBin, kBsCount };
void launch_batch_stage(int stage, int variant, const BatchVox* scenes, const BatchStage& st, int total_blocks,
                        hipStream_t s);
// scenes the batched stages handle: sides multiple of 16, fused build up to
// level 4 and every further level in the tail, at most 16384 billboards
bool batch_vox_eligible(const GridDesc& g, const VolumeDesc& vd, int S);
// launch shape of a stage for one scene: workgroups, and the variant (raster
// <P2, WORDS> and sort width) that groups scenes into one launch
void batch_stage_shape(int stage, const BatchVox& b, int* blocks, int* variant);
// segmented fill: one launch over many (pointer, count, value) segments
void launch_fill_segments(const FillSeg* d_segs, const int* d_first, int nseg, int total_blocks, hipStream_t s);
int fill_segment_blocks(uint64_t n);
// fills the kernel-derived fields of b (raster bitmap words, fused build shape)
void batch_vox_complete(BatchVox& b);

// batched cone trace: work items are global tiles in `order` (heavy first);
// waves_per_group waves per workgroup, `groups` workgroups striding over the
// items (groups = 0: one workgroup per waves_per_group items)
void launch_cone_batch(const BatchScene* scenes, const int* tile0, int nscenes, const int* order, int total_tiles,
                       int format, int cone_mode, int composite, int waves_per_group, int groups,
                       unsigned long long* counters, hipStream_t s);
// binning into a caller-given tile_count array, and the heavy-first order of
// an array of tile counts (scratch: 64 + ntiles ints)
void launch_bin_counts(const int4* sph_tiles, const int32_t* order, int S, int tiles_x, int words,
                       unsigned long long* masks, int* tile_count, hipStream_t s);
void launch_tile_order(const int* tile_count, int ntiles, int* scratch, int* tile_order, hipStream_t s);

// LDS tables of the table-addressed cone kernel: per level, word offsets of
// three arrays indexed by floor(u) + 1 for the in-volume range
// floor(u) in [-1, n - 1]: the byte offset of the clamped x corner c0, and
// the byte-offset pairs of the clamped y and z corners (c0, c1), all relative
// to the level's first texel.  words: table words; wpg: waves per workgroup.
struct ConeTab {
    uint32_t x[kMaxLevels], y[kMaxLevels], z[kMaxLevels];
    uint32_t words;
    int levels;
    int wpg;
};
// fills t for vd (format 1 or 2); false when a level's bytes reach 2^32 (the
// tables hold 32-bit byte offsets) or the format has no paired slots
bool cone_tab_plan(const VolumeDesc& vd, ConeTab& t);
size_t cone_tab_lds_bytes(const ConeTab& t);

// arguments of one cone-trace launch
struct ConeLaunch {
    GridDesc g;
    VolumeDesc vd;
    CameraView cv;
    ConeDesc cd;
    const void* levels;
    const float4* sph_world;
    const float4* sph_view;
    const int32_t* order;
    const unsigned long long* masks;
    int words;
    void* image;
    unsigned long long* counters;  // 9 counters or null
    const int* tile_order;
    MlpDev mlp;
    int learned;
    const float4* nmap;  // noise-march map (32^3 decoded RGBA8_SNORM texels) or null: no noise march
    int table_addressing = 1;  // front-to-back RGBA traces: corner offsets from per-level LDS tables
};
void launch_cone(const ConeLaunch& a, hipStream_t s);
void launch_segments(const GridDesc& g, const VolumeDesc& vd, const ConeDesc& cd, const void* levels,
                     const float* p0, int n, float* feat, float* tgt, int* valid, hipStream_t s);
void launch_occupancy(const GridDesc& g, const unsigned long long* shellbits, const unsigned long long* fillbits,
                      unsigned long long* out2, hipStream_t s);
void launch_read_state(const GridDesc& g, const unsigned long long* shellbits, const unsigned long long* fillbits,
                       uint8_t* out, hipStream_t s);
void launch_debrick(const VolumeDesc& vd, int k, const void* levels, void* out, hipStream_t s);

size_t texel_bytes(int format);

}  // namespace cv
