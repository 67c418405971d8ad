// gfx950 kernels of the billboard voxelize -> mip -> cone-trace path.
//
// Every floating-point expression is written in the operation order of the
// reference shaders (cited per kernel) and the file is compiled with
// -ffp-contract=off and IEEE division/sqrt, so FP32 results are reproducible
// bit for bit on the host.
//
// Memory layout (HBM):
//   * volume levels: 4x4x4 bricks, Morton order inside a brick (x0 y0 z0 x1 y1 z1),
//     bricks x-fastest.  RGBA texels are stored in the channel order
//     (r, a, g, b): light and density, which grey (rgba-mode) volumes need
//     alone, are one aligned 8-byte (FP32) or 4-byte (FP16) load.  A 2x2x2 trilinear footprint is one 128 B line for
//     RGBA32F texels whenever it is brick-aligned, and the 8 corner addresses
//     are sums of per-axis terms.
//   * occupancy: one 64-bit word per level-0 brick for the lit shell and one
//     for the chord fill, set with integer atomicOr (order-free, idempotent).
//   * pass-1 depth: one 64-bit key per light texel, (depth bits << 32 | draw
//     index) reduced with atomicMin: nearest wins, ties go to the first drawn
//     billboard exactly like the GL depth test LESS.
//   * camera binning: per 8x8 tile a bitmask over draw ranks, so a tile's
//     fragments are visited in draw order without a per-tile sort.

#include "cv_kernels.h"
#include "cv_mlp.h"

#include <hip/hip_runtime.h>

namespace cv {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t spread2(uint32_t v) { return (v & 1u) | ((v & 2u) << 2); }
__device__ __forceinline__ uint32_t morton6(uint32_t x, uint32_t y, uint32_t z) {
    return spread2(x) | (spread2(y) << 1) | (spread2(z) << 2);
}
// bit of (x, y, z) in a fill word: x fastest inside the 4x4x4 brick (the
// shell words and the texel bricks use Morton order)
__device__ __forceinline__ uint32_t lin6(uint32_t x, uint32_t y, uint32_t z) { return x | (y << 2) | (z << 4); }
__device__ __forceinline__ void demorton6(uint32_t m, uint32_t* x, uint32_t* y, uint32_t* z) {
    *x = (m & 1u) | ((m >> 2) & 2u);
    *y = ((m >> 1) & 1u) | ((m >> 3) & 2u);
    *z = ((m >> 2) & 1u) | ((m >> 4) & 2u);
}

// Morton offset of a clamped texel coordinate c along one axis of a level:
// the brick index times the axis' brick stride, plus c's two low bits spread
// to Morton bits (0, 3) << axis; t * 5 & 9 spreads t in [0, 3] to bits 0 and 3
__device__ __forceinline__ uint32_t morton_off_x(uint32_t c) { return ((c >> 2) << 6) | (((c & 3u) * 5u) & 9u); }
__device__ __forceinline__ uint32_t morton_off_y(uint32_t c, uint32_t srow) {
    return __umul24(c >> 2, srow) + (__umul24(c & 3u, 10u) & 18u);
}
__device__ __forceinline__ uint32_t morton_off_z(uint32_t c, uint32_t sslab) {
    return __umul24(c >> 2, sslab) + (__umul24(c & 3u, 20u) & 36u);
}

// texel (x, y, z) of a level in the bricked layout
__device__ __forceinline__ uint64_t brick_addr(const LevelDesc& L, uint32_t x, uint32_t y, uint32_t z) {
    return L.base + (uint64_t)(((z >> 2) * L.nb + (y >> 2)) * L.nb + (x >> 2)) * 64u + morton6(x & 3u, y & 3u, z & 3u);
}

__device__ __forceinline__ float h2f(uint32_t bits) {
    unsigned short s = (unsigned short)bits;
    _Float16 h;
    __builtin_memcpy(&h, &s, 2);
    return (float)h;
}
__device__ __forceinline__ uint32_t f2h(float f) {
    _Float16 h = (_Float16)f;  // v_cvt_f16_f32, round to nearest even
    unsigned short s;
    __builtin_memcpy(&s, &h, 2);
    return s;
}

// filter lerp(a, b, t) = fma(t, b - a, a): a fused multiply-add is correctly
// rounded, so the host restatement reproduces it exactly
__device__ __forceinline__ float lerpF(float a, float b, float t) { return fmaf(t, b - a, a); }

// pixel centre -> NDC: (2 * (p + 0.5)) / n - 1
__device__ __forceinline__ float pixel_ndc(int p, int n) { return (2.0f * ((float)p + 0.5f)) / (float)n - 1.0f; }

// sphereContrib (res/first_voxelize.glsl:42-48, res/conetrace_frag.glsl:177-182)
__device__ __forceinline__ float sphere_contrib(float du, float dv, float s) {
    float dist = sqrtf(du * du + dv * dv);
    float q = dist / s;
    return sqrtf(fmaxf(0.0f, 1.0f - q * q));
}

// x / range[a].  When range[a] is a power of two, 1 / range[a] is exact and
// x * (1 / range[a]) rounds the same real number as x / range[a], so the
// result is identical bit for bit (subnormals included) at a fraction of the
// cost of the IEEE division sequence.
__device__ __forceinline__ float div_range(const GridDesc& g, float x, int a) {
    return g.pow2_range ? x * g.inv_range[a] : x / g.range[a];
}
__device__ __forceinline__ float div_N(const GridDesc& g, float x) {
    return g.pow2_N ? x * g.invN : x / (float)g.N;
}

// calculateVoxelIndex with image-store bounds (res/first_voxelize.glsl:23-37)
__device__ __forceinline__ bool voxel_word_bit(const GridDesc& g, float px, float py, float pz, uint32_t* word,
                                               uint32_t* bit) {
    float fN = (float)g.N;
    float lx = fN * div_range(g, px - g.lo[0], 0);
    float ly = fN * div_range(g, py - g.lo[1], 1);
    float lz = fN * div_range(g, pz - g.lo[2], 2);
    if (!(lx > -1.0f && lx < fN && ly > -1.0f && ly < fN && lz > -1.0f && lz < fN)) return false;
    uint32_t ix = (uint32_t)(int)lx, iy = (uint32_t)(int)ly, iz = (uint32_t)(int)lz;
    *word = ((iz >> 2) * g.nb0 + (iy >> 2)) * g.nb0 + (ix >> 2);
    *bit = morton6(ix & 3u, iy & 3u, iz & 3u);
    return true;
}

// same map, returning the brick coordinates and the bit inside the fill word
// P2: the grid's extents are powers of two, x / range == x * (1 / range) exactly
template <bool P2>
__device__ __forceinline__ float div_range_t(const GridDesc& g, float x, int a) {
    return P2 ? x * g.inv_range[a] : x / g.range[a];
}
template <bool P2 = false>
__device__ __forceinline__ bool voxel_brick_bit(const GridDesc& g, float px, float py, float pz, int* bx, int* by,
                                                int* bz, uint32_t* bit) {
    float fN = (float)g.N;
    float lx, ly, lz;
    if (P2) {
        lx = fN * div_range_t<true>(g, px - g.lo[0], 0);
        ly = fN * div_range_t<true>(g, py - g.lo[1], 1);
        lz = fN * div_range_t<true>(g, pz - g.lo[2], 2);
    } else {
        lx = fN * div_range(g, px - g.lo[0], 0);
        ly = fN * div_range(g, py - g.lo[1], 1);
        lz = fN * div_range(g, pz - g.lo[2], 2);
    }
    if (!(lx > -1.0f && lx < fN && ly > -1.0f && ly < fN && lz > -1.0f && lz < fN)) return false;
    uint32_t ix = (uint32_t)(int)lx, iy = (uint32_t)(int)ly, iz = (uint32_t)(int)lz;
    *bx = (int)(ix >> 2);
    *by = (int)(iy >> 2);
    *bz = (int)(iz >> 2);
    *bit = lin6(ix & 3u, iy & 3u, iz & 3u);
    return true;
}

__device__ __forceinline__ float cv_exp2_neg(float x) {
    // 2^x for x <= 0: n = round(x), f = x - n, e^(f ln2) by Taylor degree 6
    if (!(x > -126.0f)) return 0.0f;
    float n = floorf(x + 0.5f);
    float f = x - n;
    float z = f * 0.693147180559945309f;
    float p = 1.0f / 720.0f;  // Horner with fused multiply-adds
    p = fmaf(p, z, 1.0f / 120.0f);
    p = fmaf(p, z, 1.0f / 24.0f);
    p = fmaf(p, z, 1.0f / 6.0f);
    p = fmaf(p, z, 0.5f);
    p = fmaf(p, z, 1.0f);
    p = fmaf(p, z, 1.0f);
    int ni = (int)n;
    float scale = __uint_as_float((uint32_t)(ni + 127) << 23);
    return p * scale;
}

// 1 << (b mod 64): v_lshlrev_b64 reads only the low 6 bits of its count, so
// callers pass bit indices with unmasked high bits
// bit index x | y << 2 | z << 4 of a voxel in its brick's fill word from the
// voxel coordinates, exact in the low 6 bits (all a 64-bit shift reads): two
// bit-field inserts instead of two masks, two shifts and an or
__device__ __forceinline__ uint32_t fill_bit_index(uint32_t ix, uint32_t iy, uint32_t iz) {
    const uint32_t xy = (ix & 3u) | ((iy << 2) & ~3u);
    return (xy & 15u) | ((iz << 4) & ~15u);
}
// Rotated bit layout of the pass-1 LDS words: voxel (ix, iy, iz) takes bit
// (ix + 4 iy + 16 iz) mod 64 (two shift-adds; bit64_mod keeps the low 6
// bits).  Inside brick (X, Y, Z), ix = 4 X + a etc., that is
// (lin6(a, b, c) + 4 X + 16 Y) mod 64: a bijection of the brick's voxels,
// the canonical word rotated left by (4 X + 16 Y) mod 64 (rot_to_lin6 undoes it).
__device__ __forceinline__ uint32_t rot_bit_index(uint32_t ix, uint32_t iy, uint32_t iz) {
    uint32_t t, r;  // two shift-adds (the compiler otherwise emits two shifts and an add3)
    asm("v_lshl_add_u32 %0, %1, 2, %2" : "=v"(t) : "v"(iy), "v"(ix));
    asm("v_lshl_add_u32 %0, %1, 4, %2" : "=v"(r) : "v"(iz), "v"(t));
    return r;
}
__device__ __forceinline__ unsigned long long rot_to_lin6(unsigned long long w, uint32_t X, uint32_t Y) {
    const uint32_t r = (4u * X + 16u * Y) & 63u;
    return r ? (w >> r) | (w << (64u - r)) : w;
}
// a * b + c with 24-bit unsigned factors (c and the sum mod 2^32); b and c
// uniform
__device__ __forceinline__ uint32_t mad_u24_vsv(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
    return r;
}
__device__ __forceinline__ uint32_t mad_u24_v8s(uint32_t a, uint32_t c) {
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, 8, %2" : "=v"(r) : "v"(a), "s"(c));
    return r;
}
// min of two floats that are never NaN: one v_min_f32 (fminf also quiets
// signalling NaNs, which costs a canonicalising instruction per call)
__device__ __forceinline__ float min_no_nan(float a, float b) {
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ unsigned long long bit64_mod(uint32_t b) {
    unsigned long long r;
    asm("v_lshlrev_b64 %0, %1, 1" : "=v"(r) : "v"(b));
    return r;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ void k_fill_u64(uint64_t* __restrict__ p, uint64_t n, uint64_t value) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = value;
}

// up to four fills in one launch (segments concatenated)
struct FillArgs {
    uint64_t* p[4];
    uint64_t n[4];
    uint64_t v[4];
    int nseg;
};
__global__ void k_fill_multi(FillArgs a) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t total = a.n[0] + a.n[1] + a.n[2] + a.n[3];
    for (; i < total; i += stride) {
        uint64_t r = i;
        int k = 0;
        while (k < 3 && r >= a.n[k]) r -= a.n[k++];
        a.p[k][r] = a.v[k];
    }
}

// ---------------------------------------------------------------------------
// pass 1: light-view raster (res/first_voxelize.glsl:39-64)
// ---------------------------------------------------------------------------
struct LightSphere {
    float cx, cy, cz, s;
    float vx, vy, vz;
    bool live;
};

__device__ __forceinline__ LightSphere light_sphere(const GridDesc& g, const LightView& lv, float4 b) {
    LightSphere o;
    o.cx = g.vpos[0] + b.x;
    o.cy = g.vpos[1] + b.y;
    o.cz = g.vpos[2] + b.z;
    o.s = b.w;
    o.vx = ((lv.V[0] * o.cx + lv.V[4] * o.cy) + lv.V[8] * o.cz) + lv.V[12];
    o.vy = ((lv.V[1] * o.cx + lv.V[5] * o.cy) + lv.V[9] * o.cz) + lv.V[13];
    o.vz = ((lv.V[2] * o.cx + lv.V[6] * o.cy) + lv.V[10] * o.cz) + lv.V[14];
    float dz = -o.vz;
    o.live = (o.s > 0.0f) && (dz >= lv.znear && dz <= lv.zfar);
    return o;
}

// fragment of sphere o at a light pixel: fragPos and sphereContrib;
// (hx, hy) is the pixel centre on the light's view plane
__device__ __forceinline__ bool light_fragment_at(const LightView& lv, const LightSphere& o, float hx, float hy,
                                                  float* fx, float* fy, float* fz, float* sc) {
    float du = hx - o.vx;
    float dv = hy - o.vy;
    float c = sphere_contrib(du, dv, o.s);
    *sc = c;
    if (c < 0.01f) return false;  // discard
    *fx = (o.cx + du * lv.S[0]) + dv * lv.U[0];
    *fy = (o.cy + du * lv.S[1]) + dv * lv.U[1];
    *fz = (o.cz + du * lv.S[2]) + dv * lv.U[2];
    return true;
}
__device__ __forceinline__ float light_hx(const LightView& lv, int px) { return pixel_ndc(px, lv.W) * lv.lsx + lv.lox; }
__device__ __forceinline__ float light_hy(const LightView& lv, int py) { return pixel_ndc(py, lv.H) * lv.lsy + lv.loy; }
__device__ __forceinline__ bool light_fragment(const LightView& lv, const LightSphere& o, int px, int py, float* fx,
                                               float* fy, float* fz, float* sc) {
    return light_fragment_at(lv, o, light_hx(lv, px), light_hy(lv, py), fx, fy, fz, sc);
}

// Work decomposition of pass 1: every billboard's pixel box is cut into
// kLightTile x kLightTile tiles; one workgroup rasterizes one (billboard, tile)
// item, so large spheres spread over many CUs instead of forming the tail.
// light_box: the billboard's conservative pixel box and its item count (tiles
// times the depth slices of a flat chord walk, light_slices).
__device__ __forceinline__ int light_box(const GridDesc& g, const LightView& lv, float4 board, int flat, int words,
                                         int4* out) {
    LightSphere o = light_sphere(g, lv, board);
    int4 box = make_int4(1, 1, 0, 0);
    int cnt = 0;
    if (o.live) {
        // conservative pixel box of the billboard square
        float xa = (((o.vx - o.s) - lv.lox) / lv.lsx + 1.0f) * 0.5f * (float)lv.W - 0.5f;
        float xb = (((o.vx + o.s) - lv.lox) / lv.lsx + 1.0f) * 0.5f * (float)lv.W - 0.5f;
        float ya = (((o.vy - o.s) - lv.loy) / lv.lsy + 1.0f) * 0.5f * (float)lv.H - 0.5f;
        float yb = (((o.vy + o.s) - lv.loy) / lv.lsy + 1.0f) * 0.5f * (float)lv.H - 0.5f;
        float fx0 = floorf(fminf(xa, xb)) - 1.0f, fx1 = ceilf(fmaxf(xa, xb)) + 1.0f;
        float fy0 = floorf(fminf(ya, yb)) - 1.0f, fy1 = ceilf(fmaxf(ya, yb)) + 1.0f;
        if (fx1 >= 0.0f && fy1 >= 0.0f && fx0 <= (float)(lv.W - 1) && fy0 <= (float)(lv.H - 1)) {
            box.x = (int)fmaxf(fx0, 0.0f);
            box.z = (int)fminf(fx1, (float)(lv.W - 1));
            box.y = (int)fmaxf(fy0, 0.0f);
            box.w = (int)fminf(fy1, (float)(lv.H - 1));
            cnt = ((box.z - box.x) / kLightTile + 1) * ((box.w - box.y) / kLightTile + 1);
            if (flat) cnt *= light_slices(g, lv, o.s, words);
        }
    }
    *out = box;
    return cnt;
}

// The pass-1 work list: every billboard's pixel box (lbox) and tile count,
// the exclusive scan of the counts (offsets[S] = the uncut item count) and
// the (billboard, tile) items in billboard order, cut at cap.  Three small
// kernels (256-thread workgroups): they find room on CUs that the cone trace
// of the previous frame keeps busy, where one 1024-thread workgroup waited
// for a whole CU to drain.
__device__ __forceinline__ void k_light_boxes_body(GridDesc g, LightView lv, const float4* __restrict__ boards,
                                                     int S, int4* __restrict__ lbox, int* __restrict__ counts, int chord,
                                                     int bid, int nblk) {
    const int j = bid * 256 + threadIdx.x;
    if (j >= S) return;
    int4 box;
    counts[j] = light_box(g, lv, boards[j], raster_flat(g, chord) ? 1 : 0, raster_words(g), &box);
    lbox[j] = box;
}
__global__ void __launch_bounds__(256) k_light_boxes(GridDesc g, LightView lv, const float4* __restrict__ boards,
                                                     int S, int4* __restrict__ lbox, int* __restrict__ counts, int chord) {
    k_light_boxes_body(g, lv, boards, S, lbox, counts, chord, (int)blockIdx.x, (int)gridDim.x);
}

// one workgroup: 4096 counts per round, 16 consecutive per thread
__device__ __forceinline__ void k_light_scan_body(const int* __restrict__ counts, int S, int* __restrict__ offsets) {
    __shared__ int wsum[4];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    int carry = 0;
    for (int base = 0; base < S; base += 4096) {
        int c[16];
        int tsum = 0;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int j = base + tid * 16 + e;
            c[e] = j < S ? counts[j] : 0;
            tsum += c[e];
        }
        int x = tsum;  // inclusive wave scan
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[wid] = x;
        __syncthreads();
        int wpre = 0, round = 0;
        for (int w = 0; w < 4; ++w) {
            const int ws = wsum[w];
            wpre += w < wid ? ws : 0;
            round += ws;
        }
        int o = carry + wpre + x - tsum;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int j = base + tid * 16 + e;
            if (j < S) offsets[j] = o;
            o += c[e];
        }
        carry += round;
        __syncthreads();  // wsum is rewritten next round
    }
    if (tid == 0) offsets[S] = carry;
}
__global__ void __launch_bounds__(256) k_light_scan(const int* __restrict__ counts, int S, int* __restrict__ offsets) {
    k_light_scan_body(counts, S, offsets);
}

// Depth (along the light direction, from the billboard plane) where depth
// slice m of K starts: -s + m * (2s / K), s for m = K.  The raster splits each
// chord's samples between slices at these bounds, computed the same way for
// both slices that share one, so every sample falls in exactly one slice.
__host__ __device__ __forceinline__ float slice_bound(float s, int m, int K) {
    return m >= K ? s : -s + (float)m * ((2.0f * s) / (float)K);
}

// The brick box of the chord slab behind a tile (chord fill only): the world
// box of c + du*S + dv*U + t*B, du/dv over the tile's view-plane rectangle
// (within the disc), t in [ta, tb] (the item's depth slice, [-s, s] for a
// whole chord).  Reciprocal products instead of divisions: the box keeps a
// 1-voxel margin, and the raster stores samples outside it directly, so it
// only has to be close.
__device__ void slab_box(const GridDesc& g, const LightView& lv, const LightSphere& o, float ta, float tb, int words,
                         RasterItem& it) {
    const float kW = 2.0f / (float)lv.W, kH = 2.0f / (float)lv.H;
    float du0 = fmaxf(((((float)it.x0 + 0.5f) * kW - 1.0f) * lv.lsx + lv.lox) - o.vx, -o.s);
    float du1 = fminf(((((float)it.x1 + 0.5f) * kW - 1.0f) * lv.lsx + lv.lox) - o.vx, o.s);
    float dv0 = fmaxf(((((float)it.y0 + 0.5f) * kH - 1.0f) * lv.lsy + lv.loy) - o.vy, -o.s);
    float dv1 = fminf(((((float)it.y1 + 0.5f) * kH - 1.0f) * lv.lsy + lv.loy) - o.vy, o.s);
    const float c3[3] = {o.cx, o.cy, o.cz};
    int b0[3], b1[3];
    for (int a = 0; a < 3; ++a) {
        const float kvox = (float)g.N * g.inv_range[a];
        float su0 = du0 * lv.S[a], su1 = du1 * lv.S[a], sv0 = dv0 * lv.U[a], sv1 = dv1 * lv.U[a];
        const float t0 = ta * lv.B[a], t1 = tb * lv.B[a];
        float wlo = c3[a] + fminf(su0, su1) + fminf(sv0, sv1) + fminf(t0, t1);
        float whi = c3[a] + fmaxf(su0, su1) + fmaxf(sv0, sv1) + fmaxf(t0, t1);
        float l0 = (wlo - g.lo[a]) * kvox - 1.0f;
        float l1 = (whi - g.lo[a]) * kvox + 1.0f;
        float f0 = fmaxf(floorf(l0 * 0.25f), 0.0f), f1 = fminf(floorf(l1 * 0.25f), (float)(g.nb0 - 1));
        b0[a] = (int)f0;
        b1[a] = f1 >= f0 ? (int)f1 : (int)f0 - 1;
    }
    it.bx0 = b0[0]; it.by0 = b0[1]; it.bz0 = b0[2];
    it.nbx = b1[0] - b0[0] + 1; it.nby = b1[1] - b0[1] + 1; it.nbz = b1[2] - b0[2] + 1;
    it.use_lds = du0 <= du1 && dv0 <= dv1 && it.nbx > 0 && it.nby > 0 && it.nbz > 0 &&
                 it.nbx * it.nby * it.nbz <= words;
    it.inv_x = it.use_lds ? 1.0f / (float)it.nbx : 0.0f;
    it.inv_xy = it.use_lds ? 1.0f / (float)(it.nbx * it.nby) : 0.0f;
}

// one thread per billboard enumerates its items (tiles in row order):
// (billboard, tile) into the first two words of each
__device__ __forceinline__ void k_light_expand_body(const int* __restrict__ offsets, int S,
                                                      RasterItem* __restrict__ items, int cap, int bid, int nblk) {
    const int j = bid * 256 + threadIdx.x;
    if (j >= S) return;
    const int o0 = offsets[j], c = offsets[j + 1] - o0;
    for (int t = 0; t < c && o0 + t < cap; ++t) *(int2*)&items[o0 + t] = make_int2(j, t);
}
__global__ void __launch_bounds__(256) k_light_expand(const int* __restrict__ offsets, int S,
                                                      RasterItem* __restrict__ items, int cap) {
    k_light_expand_body(offsets, S, items, cap, (int)blockIdx.x, (int)gridDim.x);
}

// one thread per item: the tile's rectangle, the billboard's light-view data
// and the slab box
__device__ __forceinline__ void k_light_prep_body(GridDesc g, LightView lv, const float4* __restrict__ boards,
                                                    const int4* __restrict__ lbox, const int* __restrict__ total_ptr,
                                                    RasterItem* __restrict__ items, int cap, int chord, int words, int bid, int nblk) {
    const int n = min(*total_ptr, cap);
    for (int i = bid * 256 + threadIdx.x; i < n; i += nblk * 256) {
    const int2 jt = *(const int2*)&items[i];
    const int j = jt.x;
    const LightSphere o = light_sphere(g, lv, boards[j]);
    // item t of the billboard: tile t / K, depth slice t % K (light_box's K)
    const int K = raster_flat(g, chord) ? light_slices(g, lv, o.s, words) : 1;
    const int t = jt.y / K, slice = jt.y - t * K;
    const int4 box = lbox[j];
    const int ntx = (box.z - box.x) / kLightTile + 1;
    RasterItem it{};
    it.j = j;
    it.slice = slice;
    it.nslices = K;
    it.x0 = box.x + (t % ntx) * kLightTile;
    it.y0 = box.y + (t / ntx) * kLightTile;
    it.x1 = min(it.x0 + kLightTile - 1, box.z);
    it.y1 = min(it.y0 + kLightTile - 1, box.w);
    it.cx = o.cx; it.cy = o.cy; it.cz = o.cz; it.s = o.s;
    it.vx = o.vx; it.vy = o.vy;
    it.inv_tw = 1.0f / (float)(it.x1 - it.x0 + 1);
    if (chord) slab_box(g, lv, o, slice_bound(o.s, slice, K), slice_bound(o.s, slice + 1, K), words, it);
    items[i] = it;
    }
}
__global__ void __launch_bounds__(256) k_light_prep(GridDesc g, LightView lv, const float4* __restrict__ boards,
                                                    const int4* __restrict__ lbox, const int* __restrict__ total_ptr,
                                                    RasterItem* __restrict__ items, int cap, int chord, int words) {
    k_light_prep_body(g, lv, boards, lbox, total_ptr, items, cap, chord, words, (int)blockIdx.x, (int)gridDim.x);
}

constexpr int kTilePixels = kLightTile * kLightTile;
constexpr int kRasterThreads = 256;  // workgroup size of the raster kernels

// P2: power-of-two extents (multiplications instead of IEEE divisions in
// the index map, and an exact k * stepSize chord counter); WORDS: size of the
// LDS chord bitmap (bigger volumes have bigger slabs behind a tile)
template <bool P2, int WORDS>
__device__ __forceinline__ void k_light_raster_body(GridDesc g, LightView lv, int chord,
                                                      const int* __restrict__ total_ptr,
                                                      const RasterItem* __restrict__ items, int cap,
                                                      unsigned long long* __restrict__ keys,
                                                      unsigned long long* __restrict__ fillbits,
                                                      unsigned long long* __restrict__ counters, int bid, int nblk) {
    // Chord stores of one tile stay inside the sphere's slab behind that
    // tile: gather them in an LDS bitmap over the slab's brick box (1-voxel
    // margin for rounding) and flush non-zero words with one global atomic
    // each.  Stores falling outside the box (none expected) go straight to HBM.
    //
    // With a power-of-two stepSize the chord loop counter i += stepSize is
    // exactly k * stepSize: covered fragments are queued in LDS with their
    // sample counts and the chords are walked from the queue, so no lane
    // idles on a pixel outside the disc.  The queue is ordered by sample
    // count, longest first (a counting sort over 128 buckets), so the 64
    // chords a wave walks together have nearly the same length.
    static_assert(WORDS <= 4096, "the flush decodes word indices below 4096 with float products");
    static_assert(kTilePixels % kRasterThreads == 0, "whole pixel rounds");
    constexpr int kRounds = kTilePixels / kRasterThreads;
    __shared__ unsigned long long lbits[WORDS];
    __shared__ float f_sx[kTilePixels], f_sy[kTilePixels], f_sz[kTilePixels];
    __shared__ int f_n[kTilePixels];
    __shared__ float h_x[kLightTile], h_y[kLightTile];  // the tile's pixel centres on the view plane
    __shared__ int nfr_s;
    __shared__ int q_start[128];  // chord-length buckets: counts, then first queue slots
    const int tid = threadIdx.x;
    const int total = min(*total_ptr, cap);
    const bool flat = P2 && chord;
    unsigned long long nfrag = 0, nchord = 0;
    // The item's setup comes from the work list.  The next item is fetched
    // while this one is rasterized: lanes 0..23 of every wave load one word
    // each (vector loads, so the LDS waits in between do not wait for them)
    // and the fields are read off those lanes at the next item; the last
    // item fetches itself again.
    if ((int)bid >= total) return;
    constexpr int kItemWords = (int)(sizeof(RasterItem) / 4);
    const int lane = tid & 63;
    int item = bid;
    int iw = lane < kItemWords ? ((const int*)(items + item))[lane] : 0;
    for (; item < total; item += nblk) {
    const int iwn = lane < kItemWords ? ((const int*)(items + min(item + (int)nblk, total - 1)))[lane] : 0;
    auto field_i = [&](int k) { return __builtin_amdgcn_readlane(iw, k); };
    auto field_f = [&](int k) { return __int_as_float(__builtin_amdgcn_readlane(iw, k)); };
    const int j = field_i(0);
    const int x0 = field_i(1), y0 = field_i(2), x1 = field_i(3), y1 = field_i(4);
    LightSphere o;
    o.cx = field_f(12); o.cy = field_f(13); o.cz = field_f(14); o.s = field_f(15);
    o.vx = field_f(16); o.vy = field_f(17);
    o.vz = 0.0f;
    o.live = true;
    int bx0 = 0, by0 = 0, bz0 = 0, nbx = 0, nby = 0, nbz = 0;
    bool use_lds = false;
    // depth slice of the item (flat walk): its share of every chord's samples;
    // slice 0 also writes the depth keys
    const int slice = field_i(21), nsl = field_i(22);
    if (chord) {
        bx0 = field_i(5); by0 = field_i(6); bz0 = field_i(7);
        nbx = field_i(8); nby = field_i(9); nbz = field_i(10);
        use_lds = field_i(11) != 0;
        if (use_lds)
            for (int w = tid; w < nbx * nby * nbz; w += kRasterThreads) lbits[w] = 0ull;
        if (tid == 0) nfr_s = 0;
        if (flat && tid < 128) q_start[tid] = 0;
    }
    if (tid < kLightTile) h_x[tid] = light_hx(lv, x0 + tid);
    else if (tid < 2 * kLightTile) h_y[tid - kLightTile] = light_hy(lv, y0 + tid - kLightTile);
    __syncthreads();
    // store one brick word of chord samples (LDS bitmap, or HBM outside it)
    auto flush = [&](int gx, int gy, int gz, unsigned long long bits) {
        if (gx < 0) return;
        if (use_lds) {
            int lx = gx - bx0, ly = gy - by0, lz = gz - bz0;
            if (lx >= 0 && lx < nbx && ly >= 0 && ly < nby && lz >= 0 && lz < nbz) {
                atomicOr(&lbits[(lz * nby + ly) * nbx + lx], bits);
                return;
            }
        }
        atomicOr(&fillbits[((uint32_t)gz * g.nb0 + (uint32_t)gy) * g.nb0 + (uint32_t)gx], bits);
    };
    // pass over the tile's pixels: depth keys; chords queued (flat) or walked
    const int tw = x1 - x0 + 1, th = y1 - y0 + 1;
    const float inv_tw = field_f(18);
    // flat: this thread's queued chords (one per pixel round) and their
    // ranks in their length bucket, placed once the buckets are counted
    float q_sx[kRounds], q_sy[kRounds], q_sz[kRounds];
    int q_n[kRounds], q_k0[kRounds], q_rank[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        q_sx[r] = q_sy[r] = q_sz[r] = 0.0f;
        q_n[r] = 0; q_k0[r] = 0; q_rank[r] = 0;
    }
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int p = tid + r * kRasterThreads;
        if (p >= tw * th) break;
        // p / tw as a float product: (p + 0.5) / tw lies at least 1 / (2 tw)
        // from an integer, far above the product's error for p < 1024
        const int ry = (int)(((float)p + 0.5f) * inv_tw), rx = p - ry * tw;
        const int px = x0 + rx, py = y0 + ry;
        float fx, fy, fz, sc;
        if (!light_fragment_at(lv, o, h_x[rx], h_y[ry], &fx, &fy, &fz, &sc)) continue;
        if (slice == 0) ++nfrag;
        float dist = o.s * sc;  // radius * sphereContrib
        if (chord) {
            // res/first_voxelize.glsl:53-58: start = fragPos - dir * dist;
            // for (float i = 0; i < 2 * dist; i += stepSize) store (0,0,0,1)
            float sx = fx - lv.B[0] * dist, sy = fy - lv.B[1] * dist, sz = fz - lv.B[2] * dist;
            float lim = 2.0f * dist;
            if (flat) {
                // k * stepSize < lim  <=>  k < lim / stepSize (exact: power-of-two scaling)
                const int n = lim > 0.0f ? (int)ceilf(lim * g.inv_step) : 0;
                // this slice's samples k in [k0, k1): the first sample index
                // whose depth -dist + k * stepSize reaches the slice bound
                int k0 = 0, k1 = n;
                if (nsl > 1) {
                    auto kb = [&](int m) {
                        if (m <= 0) return 0;
                        if (m >= nsl) return n;
                        const int k = (int)ceilf((slice_bound(o.s, m, nsl) + dist) * g.inv_step);
                        return min(max(k, 0), n);
                    };
                    k0 = kb(slice);
                    k1 = kb(slice + 1);
                }
                const int m = k1 - k0;
                if (m > 0) {
                    q_sx[r] = sx; q_sy[r] = sy; q_sz[r] = sz;
                    q_n[r] = m;
                    q_k0[r] = k0;
                    q_rank[r] = atomicAdd(&q_start[min(m, 127)], 1);
                    nchord += (unsigned long long)m;
                }
            } else {
                // consecutive samples of one chord mostly share a brick word:
                // OR them in a register and store once per word change
                int cbx = -1, cby = -1, cbz = -1;
                unsigned long long cur_b = 0ull;
                for (float i = 0.0f; i < lim; i += g.stepSize) {
                    float qx = sx + lv.B[0] * i, qy = sy + lv.B[1] * i, qz = sz + lv.B[2] * i;
                    int gx, gy, gz;
                    uint32_t b;
                    ++nchord;
                    if (!voxel_brick_bit<P2>(g, qx, qy, qz, &gx, &gy, &gz, &b)) continue;
                    if (gx != cbx || gy != cby || gz != cbz) {
                        flush(cbx, cby, cbz, cur_b);
                        cbx = gx; cby = gy; cbz = gz;
                        cur_b = 0ull;
                    }
                    cur_b |= 1ull << b;
                }
                flush(cbx, cby, cbz, cur_b);
            }
        }
        if (slice == 0) {
            float wx = fx + lv.B[0] * dist, wy = fy + lv.B[1] * dist, wz = fz + lv.B[2] * dist;
            float ex = wx - lv.np[0], ey = wy - lv.np[1], ez = wz - lv.np[2];
            float d = sqrtf((ex * ex + ey * ey) + ez * ez) / lv.clip;
            d = fminf(fmaxf(d, 0.0f), 1.0f);
            if (d < 1.0f) {
                unsigned long long key = ((unsigned long long)__float_as_uint(d) << 32) | (unsigned long long)(uint32_t)j;
                atomicMin(&keys[(size_t)py * lv.W + px], key);
            }
        }
    }
    if (flat) {
        __syncthreads();
        // bucket b's first slot = the count of the longer buckets (b' > b)
        if (tid < 64) {
            const int b0 = 127 - 2 * tid, b1 = 126 - 2 * tid;
            const int h0 = q_start[b0], h1 = q_start[b1];
            const int v = h0 + h1;
            int incl = v;
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(incl, o, 64);
                if (tid >= o) incl += t;
            }
            q_start[b0] = incl - v;
            q_start[b1] = incl - v + h0;
            if (tid == 63) nfr_s = incl;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            if (q_n[r] > 0) {
                const int slot = q_start[min(q_n[r], 127)] + q_rank[r];
                f_sx[slot] = q_sx[r]; f_sy[slot] = q_sy[r]; f_sz[slot] = q_sz[r];
                // sliced items: first sample index above the count (both < 2^16, light_slices)
                f_n[slot] = nsl > 1 ? (q_k0[r] << 16) | q_n[r] : q_n[r];
            }
        }
        __syncthreads();
        // every lane walks one queued chord: no lane idles on a pixel
        // outside the disc, and the 64 chords of a wave (consecutive slots
        // of the length-sorted queue) have nearly the same length
        const int nfr = nfr_s;
        const float fN = (float)g.N;
        const float Bx = lv.B[0], By = lv.B[1], Bz = lv.B[2];
        // N * ((q - lo) * (1 / range)): both factors are powers of two, so
        // scaling once by their product rounds identically
        const float kx = fN * g.inv_range[0], ky = fN * g.inv_range[1], kz = fN * g.inv_range[2];
        const uint32_t uN = (uint32_t)g.N;
        const uint32_t ubx = (uint32_t)bx0, uby = (uint32_t)by0, ubz = (uint32_t)bz0;
        const uint32_t unx = (uint32_t)nbx, uny = (uint32_t)nby, unz = (uint32_t)nbz;
        // word of brick (X, Y, Z) in the box: Z * unxy + Y * unx + X - wbase
        // (mod 2^32; the true value is in [0, words) for a brick in the box)
        const uint32_t unxy = unx * uny, wbase = (ubz * uny + uby) * unx + ubx;
        const float step = g.stepSize;
        // k is a power of two, so scaling the operands of each rounded step by
        // k scales its result exactly: ((s + B i) - lo) k == (s k + (B k) i) - lo k
        // bit for bit (a subnormal B i vanishes in the add in both forms)
        const float Bkx = Bx * kx, Bky = By * ky, Bkz = Bz * kz;
        const float lkx = g.lo[0] * kx, lky = g.lo[1] * ky, lkz = g.lo[2] * kz;
        for (int f0 = tid & ~63; f0 < nfr; f0 += blockDim.x) {
            // lanes past the end of the queue repeat its last chord (the
            // stores are ORs, so repeating them changes nothing): every lane
            // is active, and the inner walk's trip count is wave-uniform
            const int f = min(f0 + (tid & 63), nfr - 1);
            const float sx = f_sx[f], sy = f_sy[f], sz = f_sz[f];
            const int fn_ = f_n[f];
            const int n = nsl > 1 ? (fn_ & 0xffff) : fn_;  // samples k0 .. k0 + n - 1 of the chord
            const int k0 = nsl > 1 ? (fn_ >> 16) : 0;
            const float i0 = (float)k0 * step;  // exact: power-of-two step
            if (use_lds) {
                // the LDS word of the current brick, its bits OR-ed in a
                // register: a brick change costs one LDS atomic
                int cur_w = -1;
                unsigned long long cur_b = 0ull;
                // Each axis' voxel coordinate is monotone in k (B fixed, every
                // step rounds monotonically), so when the first and the last
                // sample lie inside the volume and the slab box, all do: such
                // chords skip the per-sample bounds tests.
                bool inner;
                {
                    const float ie = (float)(k0 + n - 1) * g.stepSize;
                    const float ax = ((sx + Bx * i0) - g.lo[0]) * kx, ay = ((sy + By * i0) - g.lo[1]) * ky;
                    const float az = ((sz + Bz * i0) - g.lo[2]) * kz;
                    const float bx = ((sx + Bx * ie) - g.lo[0]) * kx, by = ((sy + By * ie) - g.lo[1]) * ky;
                    const float bz = ((sz + Bz * ie) - g.lo[2]) * kz;
                    auto ok = [&](float l, uint32_t b0, uint32_t nb) {
                        if (!(l > -1.0f && l < fN)) return false;
                        return ((uint32_t)(int)l >> 2) - b0 < nb;
                    };
                    inner = ok(ax, ubx, unx) && ok(bx, ubx, unx) && ok(ay, uby, uny) && ok(by, uby, uny) &&
                            ok(az, ubz, unz) && ok(bz, ubz, unz);
                }
                // the longest inner chord of the wave (every lane is active here)
                int nmax = inner ? n : 0;
                for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o, 64));
                nmax = __builtin_amdgcn_readfirstlane(nmax);
                if (inner) {
                    const float skx = sx * kx, sky = sy * ky, skz = sz * kz;
                    // step k at i = min(k, n - 1) * stepSize: a lane past its
                    // last sample repeats it.  i is accumulated: with a
                    // power-of-two step every partial sum is exact.
                    const float ilast = (float)(k0 + n - 1) * step;
                    auto voxel = [&](float i, uint32_t& ix, uint32_t& iy, uint32_t& iz) {
                        ix = (uint32_t)(int)((skx + Bkx * i) - lkx);
                        iy = (uint32_t)(int)((sky + Bky * i) - lky);
                        iz = (uint32_t)(int)((skz + Bkz * i) - lkz);
                    };
                    // byte offset of the brick's word in the box,
                    // ((Z * unxy + Y * unx + X) - wbase) * 8 mod 2^32 (factors
                    // below 2^24); lin6 bit (the shift drops iz's high bits)
                    const uint32_t unxy8 = unxy * 8u, unx8 = unx * 8u, nwbase8 = 0u - wbase * 8u;
                    auto waddr = [&](uint32_t ix, uint32_t iy, uint32_t iz) {
                        return mad_u24_vsv(iz >> 2, unxy8, mad_u24_vsv(iy >> 2, unx8, mad_u24_v8s(ix >> 2, nwbase8)));
                    };
                    auto lword = [&](uint32_t a) { return (unsigned long long*)((char*)lbits + a); };
                    uint32_t ix, iy, iz;
                    voxel(i0, ix, iy, iz);
                    uint32_t cur_a = waddr(ix, iy, iz);
                    cur_b = bit64_mod(rot_bit_index(ix, iy, iz));
                    float iacc = i0;
                    for (int k = 1; k < nmax; ++k) {
                        iacc += step;
                        voxel(min_no_nan(iacc, ilast), ix, iy, iz);
                        const uint32_t a = waddr(ix, iy, iz);
                        const unsigned long long m = bit64_mod(rot_bit_index(ix, iy, iz));
                        if (a != cur_a) {
                            atomicOr(lword(cur_a), cur_b);
                            cur_a = a;
                            cur_b = 0ull;
                        }
                        cur_b |= m;
                    }
                    atomicOr(lword(cur_a), cur_b);
                    continue;
                }
                float i = i0;
                for (int k = 0; k < n; ++k, i += step) {
                    const float qx = sx + Bx * i, qy = sy + By * i, qz = sz + Bz * i;
                    const float lx = __builtin_amdgcn_fmed3f((qx - g.lo[0]) * kx, -2.0f, fN);
                    const float ly = __builtin_amdgcn_fmed3f((qy - g.lo[1]) * ky, -2.0f, fN);
                    const float lz = __builtin_amdgcn_fmed3f((qz - g.lo[2]) * kz, -2.0f, fN);
                    const uint32_t ix = (uint32_t)(int)lx, iy = (uint32_t)(int)ly, iz = (uint32_t)(int)lz;
                    const bool in = (ix | iy | iz) < uN;
                    // brick offsets in the slab's box (unsigned: below the box wraps high)
                    const uint32_t ux = (ix >> 2) - ubx, uy = (iy >> 2) - uby, uz = (iz >> 2) - ubz;
                    const bool inbox = (ux < unx) & (uy < uny) & (uz < unz);
                    const int w = (int)(__umul24(uz, unxy) + (__umul24(uy, unx) + ux));  // used when inbox
                    if (in && !inbox)  // outside the slab's box (not expected): straight to HBM
                        atomicOr(&fillbits[((iz >> 2) * g.nb0 + (iy >> 2)) * g.nb0 + (ix >> 2)],
                                 bit64_mod(fill_bit_index(ix, iy, iz)));
                    const unsigned long long m = bit64_mod(rot_bit_index(ix, iy, iz));  // LDS: rotated layout
                    const bool take = in && inbox;
                    if (take && w != cur_w) {
                        if (cur_w >= 0) atomicOr(&lbits[cur_w], cur_b);
                        cur_w = w;
                        cur_b = 0ull;
                    }
                    cur_b |= take ? m : 0ull;
                }
                if (cur_w >= 0) atomicOr(&lbits[cur_w], cur_b);
                continue;
            }
            // brick word as one key (9 bits per axis: N <= 2048 gives <= 512
            // bricks per axis), bits of the current word OR-ed in a register
            uint32_t cur_key = 0xffffffffu;
            unsigned long long cur_b = 0ull;
            for (int k = k0; k < k0 + n; ++k) {
                const float i = (float)k * g.stepSize;
                const float qx = sx + Bx * i, qy = sy + By * i, qz = sz + Bz * i;
                // voxel_brick_bit with power-of-two extents, branch free
                // clamped to [-2, N]: -1 < l < N  <=>  0 <= (int)l < N (truncation
                // toward zero), and with N a power of two the three unsigned
                // bounds fold into one test of the OR
                const float lx = __builtin_amdgcn_fmed3f((qx - g.lo[0]) * kx, -2.0f, fN);
                const float ly = __builtin_amdgcn_fmed3f((qy - g.lo[1]) * ky, -2.0f, fN);
                const float lz = __builtin_amdgcn_fmed3f((qz - g.lo[2]) * kz, -2.0f, fN);
                const uint32_t ix = (uint32_t)(int)lx, iy = (uint32_t)(int)ly, iz = (uint32_t)(int)lz;
                const bool in = (ix | iy | iz) < uN;
                // brick coordinates (< 512) in 9-bit fields
                const uint32_t key = (ix >> 2) | ((iy << 7) & 0x3fe00u) | ((iz << 16) & 0x7fc0000u);
                const unsigned long long m = 1ull << ((ix & 3u) | ((iy << 2) & 0xcu) | ((iz << 4) & 0x30u));
                if (in && key != cur_key) {
                    if (cur_key != 0xffffffffu)
                        flush((int)(cur_key & 511u), (int)((cur_key >> 9) & 511u), (int)(cur_key >> 18), cur_b);
                    cur_key = key;
                    cur_b = 0ull;
                }
                cur_b |= in ? m : 0ull;
            }
            if (cur_key != 0xffffffffu)
                flush((int)(cur_key & 511u), (int)((cur_key >> 9) & 511u), (int)(cur_key >> 18), cur_b);
        }
    }
    if (chord) {
        __syncthreads();
        if (use_lds) {
            // w / nbx and w / (nbx nby) as float products: exact for w < 4096
            // (the error stays below the 1 / (2 n) margin of the floor)
            const float inv_x = field_f(19), inv_xy = field_f(20);
            for (int w = tid; w < nbx * nby * nbz; w += blockDim.x) {
                unsigned long long bits = lbits[w];
                if (!bits) continue;
                const float fw = (float)w + 0.5f;
                const int q = (int)(fw * inv_x), lz = (int)(fw * inv_xy);
                const int lx = w - q * nbx, ly = q - lz * nby;
                uint32_t gw = ((uint32_t)(lz + bz0) * g.nb0 + (uint32_t)(ly + by0)) * g.nb0 + (uint32_t)(lx + bx0);
                // the flat walk's words are in the rotated layout
                if (flat) bits = rot_to_lin6(bits, (uint32_t)(lx + bx0), (uint32_t)(ly + by0));
                atomicOr(&fillbits[gw], bits);
            }
        }
    }
    __syncthreads();  // the next item reuses the bitmap, the queue and the pixel centres
    iw = iwn;
    }  // items
    if (counters) {
        nfrag = wave_sum_u64(nfrag);
        nchord = wave_sum_u64(nchord);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&counters[0], nfrag);
            atomicAdd(&counters[1], nchord);
        }
    }
}
template <bool P2, int WORDS>
__global__ void __launch_bounds__(256) k_light_raster(GridDesc g, LightView lv, int chord,
                                                      const int* __restrict__ total_ptr,
                                                      const RasterItem* __restrict__ items, int cap,
                                                      unsigned long long* __restrict__ keys,
                                                      unsigned long long* __restrict__ fillbits,
                                                      unsigned long long* __restrict__ counters) {
    k_light_raster_body<P2, WORDS>(g, lv, chord, total_ptr, items, cap, keys, fillbits, counters, (int)blockIdx.x, (int)gridDim.x);
}

// ---------------------------------------------------------------------------
// depth resolve + pass 2 (res/second_voxelize.glsl:34-51)
// ---------------------------------------------------------------------------
// posmap (optional): the pass-1 position map (RGBA32F, 0 where no fragment
// won); shellbits (optional): the 9-voxel splat of pass 2.  A frame runs the
// splat alone; the position map is rebuilt from the same keys on request.
// A workgroup covers 16 x 16 light texels; neighbouring texels hit the same
// shell words (the light map samples each voxel a few times), so the words
// are OR-ed in an LDS hash table and flushed with one global atomic each.
constexpr int kSplatSlots = 512;
constexpr uint32_t kSplatEmpty = 0xffffffffu;

__device__ __forceinline__ void k_light_splat_body(GridDesc g, LightView lv, const float4* __restrict__ boards,
                                                     const unsigned long long* __restrict__ keys,
                                                     float4* __restrict__ posmap,
                                                     unsigned long long* __restrict__ shellbits,
                                                     unsigned long long* __restrict__ counters, int bid, int nblk) {
    __shared__ uint32_t hkey[kSplatSlots];
    __shared__ unsigned long long hbits[kSplatSlots];
    const int t = threadIdx.x;
    if (shellbits) {
        for (int i = t; i < kSplatSlots; i += 256) { hkey[i] = kSplatEmpty; hbits[i] = 0ull; }
        __syncthreads();
    }
    const int tiles_x = (lv.W + 15) / 16;
    const int px = (int)(bid % tiles_x) * 16 + (t & 15), py = (int)(bid / tiles_x) * 16 + (t >> 4);
    const bool inside = px < lv.W && py < lv.H;
    const size_t pix = (size_t)py * lv.W + px;
    const unsigned long long key = inside ? keys[pix] : ~0ull;
    if (inside && key == ~0ull && posmap) posmap[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (key != ~0ull) {
        int j = (int)(uint32_t)(key & 0xffffffffull);
        LightSphere o = light_sphere(g, lv, boards[j]);
        float fx, fy, fz, sc;
        light_fragment(lv, o, px, py, &fx, &fy, &fz, &sc);
        float dist = o.s * sc;
        float wx = fx + lv.B[0] * dist, wy = fy + lv.B[1] * dist, wz = fz + lv.B[2] * dist;
        if (posmap) posmap[pix] = make_float4(wx, wy, wz, 1.0f);
        if (counters) atomicAdd(&counters[2], 1ull);
        if (shellbits) {
            // stepSize * normalize(vec3(+-1,+-1,+-1)): 1/sqrt(3) rounded once to FP32
            const float k = 1.0f / sqrtf(3.0f);
            const float off = g.stepSize * k;
            // the 9 stores lie within 2 voxels per axis: merge those sharing a
            // brick word in registers first
            uint32_t ws[9];
            unsigned long long bs[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                float qx = wx, qy = wy, qz = wz;
                if (q > 0) {
                    const int u = q - 1;
                    qx = wx + ((u & 4) ? -off : off);
                    qy = wy + ((u & 2) ? -off : off);
                    qz = wz + ((u & 1) ? -off : off);
                }
                uint32_t w, b;
                bool ok = voxel_word_bit(g, qx, qy, qz, &w, &b);
                ws[q] = ok ? w : kSplatEmpty;
                bs[q] = ok ? (1ull << b) : 0ull;
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                bool first = ws[q] != kSplatEmpty;
#pragma unroll
                for (int u = 0; u < q; ++u) first = first && ws[u] != ws[q];
                unsigned long long bits = bs[q];
#pragma unroll
                for (int u = q + 1; u < 9; ++u) bits |= ws[u] == ws[q] ? bs[u] : 0ull;
                if (!first) continue;
                // LDS hash insert (linear probing, bounded); overflow goes to HBM
                uint32_t h = (ws[q] * 2654435761u) >> 23;  // 9 bits
                bool done = false;
                for (int probe = 0; probe < 16 && !done; ++probe) {
                    uint32_t prev = atomicCAS(&hkey[h], kSplatEmpty, ws[q]);
                    if (prev == kSplatEmpty || prev == ws[q]) {
                        atomicOr(&hbits[h], bits);
                        done = true;
                    } else {
                        h = (h + 1u) & (uint32_t)(kSplatSlots - 1);
                    }
                }
                if (!done) atomicOr(&shellbits[ws[q]], bits);
            }
        }
    }
    if (shellbits) {
        __syncthreads();
        for (int i = t; i < kSplatSlots; i += 256)
            if (hkey[i] != kSplatEmpty) atomicOr(&shellbits[hkey[i]], hbits[i]);
    }
}
__global__ void __launch_bounds__(256) k_light_splat(GridDesc g, LightView lv, const float4* __restrict__ boards,
                                                     const unsigned long long* __restrict__ keys,
                                                     float4* __restrict__ posmap,
                                                     unsigned long long* __restrict__ shellbits,
                                                     unsigned long long* __restrict__ counters) {
    k_light_splat_body(g, lv, boards, keys, posmap, shellbits, counters, (int)blockIdx.x, (int)gridDim.x);
}

// ---------------------------------------------------------------------------
// level 0 from the occupancy bitmaps; value noise density modulation
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lattice_hash(uint32_t seed, int ix, int iy, int iz) {
    uint32_t h = seed;
    h ^= (uint32_t)ix * 0x8da6b343u;
    h ^= (uint32_t)iy * 0xd8163841u;
    h ^= (uint32_t)iz * 0xcb1ab31fu;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
__device__ __forceinline__ float lattice_value(uint32_t seed, int ix, int iy, int iz) {
    return (float)(lattice_hash(seed, ix, iy, iz) >> 8) * (1.0f / 16777216.0f);
}
__device__ __forceinline__ float smoothw(float f) { return (f * f) * (3.0f - 2.0f * f); }
__device__ __forceinline__ float lerpf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// lattice values at the 8 corners of cell (ix, iy, iz), corner k at
// (ix + (k & 1), iy + ((k >> 1) & 1), iz + (k >> 2))
__device__ __forceinline__ void cell_corners(uint32_t seed, int ix, int iy, int iz, float c[8]) {
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = lattice_value(seed, ix + (k & 1), iy + ((k >> 1) & 1), iz + (k >> 2));
}
// cell_corners for a cell every active lane of the wave shares, the hash on
// scalar registers (each value passes through readfirstlane, which keeps the
// compiler from merging this hash with the per-lane one)
__device__ __forceinline__ void cell_corners_wave(uint32_t seed, int ix, int iy, int iz, float c[8]) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t h = (uint32_t)__builtin_amdgcn_readfirstlane(
            (int)lattice_hash(seed, ix + (k & 1), iy + ((k >> 1) & 1), iz + (k >> 2)));
        c[k] = (float)(h >> 8) * (1.0f / 16777216.0f);
    }
}
// smoothstep-weighted trilinear blend of the corners: x, then y, then z
__device__ __forceinline__ float cell_blend(const float c[8], float wx, float wy, float wz) {
    float x00 = lerpf(c[0], c[1], wx), x10 = lerpf(c[2], c[3], wx);
    float x01 = lerpf(c[4], c[5], wx), x11 = lerpf(c[6], c[7], wx);
    float y0 = lerpf(x00, x10, wy), y1 = lerpf(x01, x11, wy);
    return lerpf(y0, y1, wz);
}

__device__ float value_noise(uint32_t seed, float px, float py, float pz) {
    float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    float wx = smoothw(px - fx), wy = smoothw(py - fy), wz = smoothw(pz - fz);
    float c[8];
    cell_corners(seed, ix, iy, iz, c);
    return cell_blend(c, wx, wy, wz);
}

__device__ float fbm_voxel(const VolumeDesc& vd, int N, int x, int y, int z) {
    float fN = (float)N;
    float u = ((float)x + 0.5f) / fN, w = ((float)y + 0.5f) / fN, t = ((float)z + 0.5f) / fN;
    float freq = vd.noise_base_freq, amp = 1.0f, sum = 0.0f, norm = 0.0f;
    for (int o = 0; o < vd.noise_octaves; ++o) {
        uint32_t seed = vd.noise_seed + (uint32_t)o * 0x9e3779b9u;
        sum = sum + amp * value_noise(seed, u * freq, w * freq, t * freq);
        norm = norm + amp;
        freq = freq * vd.noise_lacunarity;
        amp = amp * vd.noise_gain;
    }
    return sum / norm;
}

// fbm_voxel of the 2x2x2 block at even (x0, y0, z0), voxel v at
// (x0 + (v & 1), y0 + ((v >> 1) & 1), z0 + (v >> 2)); every voxel's value is
// computed (a caller ignores those it does not need).  The per-voxel arithmetic
// is that of fbm_voxel, octave by octave, so the results are identical.  When
// the block lies in one lattice cell (always, for cells of two voxels or more
// aligned to even coordinates) the corners are hashed once, the smoothstep
// weights once per axis coordinate and the x and y lerps once per distinct
// (x) and (x, y) pair; when every active lane of the wave is in the same cell
// the hash inputs are wave-uniform and the hash runs on scalar registers.
__device__ void fbm_block8(const VolumeDesc& vd, int N, int x0, int y0, int z0, float out[8]) {
    const float fN = (float)N;
    float u[2], w[2], t[2];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        u[d] = ((float)(x0 + d) + 0.5f) / fN;
        w[d] = ((float)(y0 + d) + 0.5f) / fN;
        t[d] = ((float)(z0 + d) + 0.5f) / fN;
    }
    float sum[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) sum[v] = 0.0f;
    float freq = vd.noise_base_freq, amp = 1.0f, norm = 0.0f;
    for (int o = 0; o < vd.noise_octaves; ++o) {
        const uint32_t seed = vd.noise_seed + (uint32_t)o * 0x9e3779b9u;
        float px[2], py[2], pz[2], fx[2], fy[2], fz[2];
        int ix[2], iy[2], iz[2];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            px[d] = u[d] * freq; py[d] = w[d] * freq; pz[d] = t[d] * freq;
            fx[d] = floorf(px[d]); fy[d] = floorf(py[d]); fz[d] = floorf(pz[d]);
            ix[d] = (int)fx[d]; iy[d] = (int)fy[d]; iz[d] = (int)fz[d];
        }
        float wx[2], wy[2], wz[2];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            wx[d] = smoothw(px[d] - fx[d]); wy[d] = smoothw(py[d] - fy[d]); wz[d] = smoothw(pz[d] - fz[d]);
        }
        if (ix[1] == ix[0] && iy[1] == iy[0] && iz[1] == iz[0]) {
            float c[8];
            const int sx = __builtin_amdgcn_readfirstlane(ix[0]);
            const int sy = __builtin_amdgcn_readfirstlane(iy[0]);
            const int sz = __builtin_amdgcn_readfirstlane(iz[0]);
            if (__builtin_amdgcn_ballot_w64(ix[0] != sx || iy[0] != sy || iz[0] != sz) == 0ull)
                cell_corners_wave(seed, sx, sy, sz, c);
            else
                cell_corners(seed, ix[0], iy[0], iz[0], c);
            // cell_blend's x lerps per x coordinate, y lerps per (x, y)
            float x00[2], x10[2], x01[2], x11[2], y0[4], y1[4];
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                x00[dx] = lerpf(c[0], c[1], wx[dx]); x10[dx] = lerpf(c[2], c[3], wx[dx]);
                x01[dx] = lerpf(c[4], c[5], wx[dx]); x11[dx] = lerpf(c[6], c[7], wx[dx]);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int dx = k & 1, dy = k >> 1;
                y0[k] = lerpf(x00[dx], x10[dx], wy[dy]);
                y1[k] = lerpf(x01[dx], x11[dx], wy[dy]);
            }
#pragma unroll
            for (int v = 0; v < 8; ++v) sum[v] = sum[v] + amp * lerpf(y0[v & 3], y1[v & 3], wz[v >> 2]);
        } else {
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const int dx = v & 1, dy = (v >> 1) & 1, dz = v >> 2;
                float e[8];
                cell_corners(seed, ix[dx], iy[dy], iz[dz], e);
                sum[v] = sum[v] + amp * cell_blend(e, wx[dx], wy[dy], wz[dz]);
            }
        }
        norm = norm + amp;
        freq = freq * vd.noise_lacunarity;
        amp = amp * vd.noise_gain;
    }
#pragma unroll
    for (int v = 0; v < 8; ++v) out[v] = sum[v] / norm;
}

// ---------------------------------------------------------------------------
// level 0 and the mip chain
//
// RGBA levels hold (r, a, r+, a+) slots (x+ = min(x + 1, n - 1)).  Every
// writer below fills both halves of its slots, except the fused build at a
// region's last column: there the second half is written by the next region
// (from its own first column), the first half by the region itself.
// ---------------------------------------------------------------------------
// (r, a) of a level-0 voxel; noise: its fbm value (used when the volume has
// noise and a != 0); RGBA16F values are rounded to half
template <int FMT>
__device__ __forceinline__ float2 level0_ra(const VolumeDesc& vd, bool shell, bool fill, float noise) {
    float r, a;
    if (vd.mode == 0) {
        r = a = shell ? 1.0f : 0.0f;
    } else {
        if (shell) { r = a = 1.0f; }
        else if (fill) { r = 0.0f; a = 1.0f; }
        else { r = a = 0.0f; }
        if (vd.noise_octaves > 0 && a != 0.0f) a = a * noise;
    }
    if (FMT == 2) return make_float2(h2f(f2h(r)), h2f(f2h(a)));
    return make_float2(r, a);
}

// parent of 8 children summed in order (RGBA16F rounded to half)
template <int FMT>
__device__ __forceinline__ float2 mip_parent2(float2 s, float cnt) {
    float2 r = make_float2(s.x / cnt, s.y / cnt);
    if (FMT == 2) r = make_float2(h2f(f2h(r.x)), h2f(f2h(r.y)));
    return r;
}
__device__ __forceinline__ float2 add2(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }

// slot stores and loads of the RGBA formats (levels indexed in texels)
template <int FMT>
__device__ __forceinline__ void store_pair(void* levels, uint64_t idx, float2 own, float2 nb) {
    if (FMT == 1) {
        ((float4*)levels)[idx] = make_float4(own.x, own.y, nb.x, nb.y);
    } else {
        ((uint2*)levels)[idx] = make_uint2(f2h(own.x) | (f2h(own.y) << 16), f2h(nb.x) | (f2h(nb.y) << 16));
    }
}
template <int FMT>
__device__ __forceinline__ void store_half(void* levels, uint64_t idx, int half, float2 v) {
    if (FMT == 1) ((float2*)levels)[2 * idx + half] = v;
    else ((uint32_t*)levels)[2 * idx + half] = f2h(v.x) | (f2h(v.y) << 16);
}
__device__ __forceinline__ float2 load_ra(const void* levels, int format, uint64_t idx) {
    if (format == 1) return ((const float2*)levels)[2 * idx];
    const uint32_t u = ((const uint32_t*)levels)[2 * idx];
    return make_float2(h2f(u & 0xffffu), h2f(u >> 16));
}
__device__ __forceinline__ void store_pair_rt(void* levels, int format, uint64_t idx, float2 own, float2 nb) {
    if (format == 1) store_pair<1>(levels, idx, own, nb);
    else store_pair<2>(levels, idx, own, nb);
}

// one wavefront per level-0 brick; lane = Morton index inside the brick
// (sides that are not a multiple of 16; the fused build covers the others)
__global__ void __launch_bounds__(256) k_level0(GridDesc g, VolumeDesc vd, const unsigned long long* __restrict__ shellbits,
                                                const unsigned long long* __restrict__ fillbits, void* __restrict__ levels,
                                                uint32_t nbricks) {
    const uint32_t brick = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t m = threadIdx.x & 63u;
    if (brick >= nbricks) return;  // whole waves
    const unsigned long long sh = shellbits[brick];
    const unsigned long long fl = fillbits ? fillbits[brick] : 0ull;
    const uint32_t nb = g.nb0;
    const uint32_t bx = brick % nb, by = (brick / nb) % nb, bz = brick / (nb * nb);
    uint32_t lx, ly, lz;
    demorton6(m, &lx, &ly, &lz);
    const int x = (int)(bx * 4u + lx), y = (int)(by * 4u + ly), z = (int)(bz * 4u + lz);
    const bool inside = x < g.N && y < g.N && z < g.N;
    const bool shell = inside && ((sh >> m) & 1ull);
    const bool fill = inside && ((fl >> lin6(lx, ly, lz)) & 1ull);
    const uint64_t idx = vd.lv[0].base + (uint64_t)brick * 64u + m;
    if (vd.format == 0) {
        ((uint8_t*)levels)[idx] = shell ? (uint8_t)255 : (uint8_t)0;
        return;
    }
    const bool noisy = vd.mode != 0 && vd.noise_octaves > 0;
    float2 v = vd.format == 1 ? level0_ra<1>(vd, shell, fill, 1.0f) : level0_ra<2>(vd, shell, fill, 1.0f);
    if (noisy && v.y != 0.0f) {
        const float nz = fbm_voxel(vd, g.N, x, y, z);
        v = vd.format == 1 ? level0_ra<1>(vd, shell, fill, nz) : level0_ra<2>(vd, shell, fill, nz);
    }
    // the +x neighbour: the lane of (lx + 1) in this brick, the next brick's
    // first column for lx = 3, the voxel itself at x = N - 1
    const int src = (int)morton6(min(lx + 1u, 3u), ly, lz);
    float2 nbv = make_float2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
    if (!inside) {
        v = nbv = make_float2(0.0f, 0.0f);
    } else if (x + 1 >= g.N) {
        nbv = v;
    } else if (lx == 3u) {
        const uint32_t nbrick = brick + 1u;
        const uint32_t mm = morton6(0u, ly, lz);
        const bool nsh = (shellbits[nbrick] >> mm) & 1ull;
        const bool nfi = fillbits ? ((fillbits[nbrick] >> lin6(0u, ly, lz)) & 1ull) != 0 : false;
        float2 w = vd.format == 1 ? level0_ra<1>(vd, nsh, nfi, 1.0f) : level0_ra<2>(vd, nsh, nfi, 1.0f);
        if (noisy && w.y != 0.0f) {
            const float nz = fbm_voxel(vd, g.N, x + 1, y, z);
            w = vd.format == 1 ? level0_ra<1>(vd, nsh, nfi, nz) : level0_ra<2>(vd, nsh, nfi, nz);
        }
        nbv = w;
    }
    store_pair_rt(levels, vd.format, idx, v, nbv);
}

// The 8 noise texels of one 2x2x2 block (texture index t0 .. t0 + 7, t0 a
// multiple of 8): binary16 for RGBA16F volumes, float for RGBA32F.
template <int FMT>
struct NoiseRaw;
template <>
struct NoiseRaw<2> {
    uint4 h;
    __device__ __forceinline__ void load(const void* tex, uint64_t t0) { h = ((const uint4*)tex)[t0 / 8u]; }
    __device__ __forceinline__ void unpack(float (&nz)[8]) const {
        const uint32_t hw[4] = {h.x, h.y, h.z, h.w};
#pragma unroll
        for (int c = 0; c < 8; ++c) nz[c] = h2f((hw[c >> 1] >> (16 * (c & 1))) & 0xffffu);
    }
};
template <>
struct NoiseRaw<1> {
    float4 a, b;
    __device__ __forceinline__ void load(const void* tex, uint64_t t0) {
        a = ((const float4*)tex)[t0 / 4u];
        b = ((const float4*)tex)[t0 / 4u + 1u];
    }
    __device__ __forceinline__ void unpack(float (&nz)[8]) const {
        nz[0] = a.x; nz[1] = a.y; nz[2] = a.z; nz[3] = a.w;
        nz[4] = b.x; nz[5] = b.y; nz[6] = b.z; nz[7] = b.w;
    }
};

// A level-0 texel staged in LDS for the slot stores: RGBA16F as the binary16
// bits of (r, a) in one word (r in the low half), RGBA32F as floats.
template <int FMT>
struct Staged {
    using T = float2;
    static __device__ __forceinline__ T pack(float2 v) { return v; }
    static __device__ __forceinline__ void pair(void* levels, uint64_t idx, T own, T nb) {
        ((float4*)levels)[idx] = make_float4(own.x, own.y, nb.x, nb.y);
    }
    static __device__ __forceinline__ void half(void* levels, uint64_t idx, int h, T v) {
        ((float2*)levels)[2 * idx + h] = v;
    }
};
template <>
struct Staged<2> {
    using T = uint32_t;
    static __device__ __forceinline__ T pack(float2 v) { return f2h(v.x) | (f2h(v.y) << 16); }
    static __device__ __forceinline__ void pair(void* levels, uint64_t idx, T own, T nb) {
        ((uint2*)levels)[idx] = make_uint2(own, nb);
    }
    static __device__ __forceinline__ void half(void* levels, uint64_t idx, int h, T v) {
        ((uint32_t*)levels)[2 * idx + h] = v;
    }
};

// ---------------------------------------------------------------------------
// fused level 0 + mips 1..4 for one 16^3 region of level 0 (N % 16 == 0)
//
// Level-0 texel values come from the occupancy words (plus noise); the 2x2x2
// children of a level-1 texel are the 8 consecutive Morton texels of one
// level-0 brick (Morton bit order x0 y0 z0 is the z, y, x summation order of
// the mip rule), so a thread owning 8 consecutive texels forms its level-1
// texel in registers.  Levels 2..4 are reduced from LDS.  Every level's value
// is rounded to the storage format before it feeds the next level, exactly
// like a level-by-level build.  Level 0 is staged in LDS half a region at a
// time (16 x 16 x 8 texels), which gives the +x neighbours of all but the
// region's last column; those are this region's first column, which it
// writes into the previous region's last slots (second halves only).
// ---------------------------------------------------------------------------
//
// Every value a region's workgroup writes (its texels of levels 0..4, and the
// second halves of the previous region's last column) derives from the
// region's own voxels, and an empty voxel is (0, 0) at every level.  With
// zflags, a region whose occupancy words are all zero and whose flag says
// that its slots already hold zero returns at once; otherwise it writes and
// records whether it wrote zeros.
template <int FMT, bool TEX>
__device__ __forceinline__ void k_build_fused_body(GridDesc g, VolumeDesc vd,
                                                     const unsigned long long* __restrict__ shellbits,
                                                     const unsigned long long* __restrict__ fillbits,
                                                     void* __restrict__ levels, uint8_t* __restrict__ zflags, int nlev,
                                                     uint32_t rpa, int bid, int nblk) {
    static_assert(FMT == 1 || FMT == 2, "RGBA formats");
    using ST = Staged<FMT>;
    __shared__ typename ST::T stage[2048];  // [brick of the half region][Morton]
    __shared__ float2 l1[512];
    __shared__ float2 l2[64];
    __shared__ float2 l3[8];
    __shared__ unsigned long long s_words[128];  // shell words of the region's 64 bricks, then fill words
    const uint32_t tid = threadIdx.x;
    const uint32_t r = bid;
    const uint32_t rx = r % rpa, ry = (r / rpa) % rpa, rz = r / (rpa * rpa);
    const bool last_x = rx + 1u == rpa;
    const uint32_t nb0 = g.nb0;
    // threads 0..63 load the shell words of the region's 64 bricks, 64..127
    // the fill words (brick lb = x + 4 y + 16 z inside the region)
    unsigned long long w = 0ull;
    if (tid < 128u) {
        const uint32_t lb = tid & 63u;
        const uint32_t brick = ((4u * rz + (lb >> 4)) * nb0 + 4u * ry + ((lb >> 2) & 3u)) * nb0 + 4u * rx + (lb & 3u);
        if (tid < 64u) w = shellbits[brick];
        else if (fillbits) w = fillbits[brick];
        s_words[tid] = w;
    }
    if (zflags) {
        const bool stored_zero = zflags[r] != 0;
        const bool any = __syncthreads_or(w != 0ull) != 0;
        if (!any && stored_zero) return;
        if (tid == 0) zflags[r] = any ? 0 : 1;
    } else {
        __syncthreads();
    }
    const bool noisy = vd.mode != 0 && vd.noise_octaves > 0;
    // Each thread owns one 2x2x2 block of texels per half region: Morton
    // slots q*8 .. q*8+7 of brick lb, i.e. the block at brick offset
    // 2 * (q & 1, (q >> 1) & 1, q >> 2).  Its occupancy and its noise texels
    // (the texture holds the values fbm_block8 computes) are fetched for both
    // halves before either is built, so the two texture reads overlap.
    // Texel c of a block (c = dx + 2 dy + 4 dz, the Morton order) has shell
    // bit q*8 + c of the brick's shell word and fill bit lin6(ox + dx, oy + dy,
    // oz + dz) = base + dx + 4 dy + 16 dz of its fill word: both gathered into
    // one byte each, bit c.
    uint32_t shb2[2], fb2[2];
    NoiseRaw<FMT> nraw[2];
#pragma unroll
    for (uint32_t rep = 0; rep < 2; ++rep) {
        const uint32_t grp = tid + rep * 256u;
        const uint32_t lb = grp >> 3, q = grp & 7u;
        const unsigned long long sh = s_words[lb], fl = s_words[64u + lb];
        const uint32_t ox = 2u * (q & 1u), oy = 2u * ((q >> 1) & 1u), oz = 2u * (q >> 2);
        const uint32_t fw = (uint32_t)(fl >> (ox + 4u * oy + 16u * oz));
        shb2[rep] = (uint32_t)(sh >> (q * 8u)) & 0xffu;
        fb2[rep] = (fw & 3u) | ((fw >> 2) & 0xcu) | ((fw >> 12) & 0x30u) | ((fw >> 14) & 0xc0u);
        const uint32_t occ = shb2[rep] | fb2[rep];
        nraw[rep] = NoiseRaw<FMT>{};
        if (TEX && noisy && occ) {
            const uint32_t brick = ((4u * rz + (lb >> 4)) * nb0 + 4u * ry + ((lb >> 2) & 3u)) * nb0 + 4u * rx + (lb & 3u);
            nraw[rep].load(vd.ntex, (uint64_t)brick * 64u + q * 8u);
        }
    }
#pragma unroll
    for (uint32_t rep = 0; rep < 2; ++rep) {
        const uint32_t grp = tid + rep * 256u;
        const uint32_t lb = grp >> 3, q = grp & 7u;
        const uint32_t lbx = lb & 3u, lby = (lb >> 2) & 3u, lbz = lb >> 4;
        const uint32_t gbx = 4u * rx + lbx, gby = 4u * ry + lby, gbz = 4u * rz + lbz;
        const uint32_t ox = 2u * (q & 1u), oy = 2u * ((q >> 1) & 1u), oz = 2u * (q >> 2);
        const uint32_t shb = shb2[rep], fb = fb2[rep];
        float2 acc = make_float2(0.0f, 0.0f);
        if constexpr (FMT == 2 && TEX) {
            // binary16 bits straight from the texture: r and a are 0, 1 or a
            // texture value, all exact halves, so level0_ra's rounding to half
            // is the identity
            const uint32_t hw[4] = {nraw[rep].h.x, nraw[rep].h.y, nraw[rep].h.z, nraw[rep].h.w};
#pragma unroll
            for (uint32_t c = 0; c < 8; ++c) {
                const bool shell = (shb >> c) & 1u, fill = (fb >> c) & 1u;
                const bool on = vd.mode == 0 ? shell : (shell || fill);
                const uint32_t nbits = (hw[c >> 1] >> (16u * (c & 1u))) & 0xffffu;
                const uint32_t ab = on ? (noisy ? nbits : 0x3c00u) : 0u;
                stage[tid * 8u + c] = (shell ? 0x3c00u : 0u) | (ab << 16);
                acc = add2(acc, make_float2(shell ? 1.0f : 0.0f, h2f(ab)));
            }
        } else {
            float nz[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) nz[c] = 1.0f;
            if (noisy && (shb | fb)) {
                if (TEX) nraw[rep].unpack(nz);
                else fbm_block8(vd, g.N, (int)(gbx * 4u + ox), (int)(gby * 4u + oy), (int)(gbz * 4u + oz), nz);
            }
#pragma unroll
            for (uint32_t c = 0; c < 8; ++c) {
                const float2 v = level0_ra<FMT>(vd, (shb >> c) & 1u, (fb >> c) & 1u, nz[c]);
                stage[tid * 8u + c] = ST::pack(v);
                acc = add2(acc, v);
            }
        }
        if (nlev > 1) {
            const uint32_t x1 = 2u * lbx + (q & 1u), y1 = 2u * lby + ((q >> 1) & 1u), z1 = 2u * lbz + (q >> 2);
            l1[(z1 * 8u + y1) * 8u + x1] = mip_parent2<FMT>(acc, 8.0f);
        }
        __syncthreads();
        if (rx > 0u && tid < 128u) {
            // first column (x = 0) -> second halves of the previous region's x = 15
            const uint32_t ey = tid & 15u, ez = tid >> 4;
            const uint32_t sbl = ((ez >> 2) << 4) | ((ey >> 2) << 2);
            const uint32_t pbrick = ((4u * rz + 2u * rep + (ez >> 2)) * nb0 + 4u * ry + (ey >> 2)) * nb0 + 4u * rx - 1u;
            ST::half(levels, vd.lv[0].base + (uint64_t)pbrick * 64u + morton6(3u, ey & 3u, ez & 3u), 1,
                     stage[(sbl << 6) | morton6(0u, ey & 3u, ez & 3u)]);
        }
        // 2048 staged texels = 32 bricks of this half region, 64 per store
        // instruction (lane = Morton slot: coalesced).  Texel i = tid + 256 k
        // is slot sm = tid & 63 of staged brick w4 + 4 k (w4 = tid >> 6): the
        // slot, its x column and so the neighbour's slot are the same for all
        // k, and the brick advances by one row (y) or one z-layer.
        {
            const uint32_t sm = tid & 63u, w4 = tid >> 6;  // w4 = the staged brick's x (0..3)
            uint32_t tx, ty, tz;
            demorton6(sm, &tx, &ty, &tz);
            // +x neighbour: the same staged brick, the next one in x, the voxel
            // itself (the volume's last column), or none here (the region's
            // last column: the next region writes that half)
            const bool edge = tx == 3u && w4 == 3u;
            const bool half = edge && !last_x;
            const uint32_t nb = edge ? sm : morton6(tx < 3u ? tx + 1u : 0u, ty, tz) + (tx == 3u ? 64u : 0u);
            const uint64_t b0 = ((uint64_t)(4u * rz + 2u * rep) * nb0 + 4u * ry) * nb0 + 4u * rx + w4;
            // slot offsets of the next brick row (y) and layer (z): uniform
            const uint64_t rowS = (uint64_t)nb0 * 64u, zS = (uint64_t)nb0 * nb0 * 64u;
            const uint64_t slot0 = vd.lv[0].base + b0 * 64u + sm;
            typename ST::T own[8], nbv[8];
#pragma unroll
            for (uint32_t k = 0; k < 8u; ++k) {
                own[k] = stage[tid + 256u * k];
                nbv[k] = stage[((w4 + 4u * k) << 6) + nb];
            }
            if (half) {
#pragma unroll
                for (uint32_t k = 0; k < 8u; ++k) ST::half(levels, slot0 + (k & 3u) * rowS + (k >> 2) * zS, 0, own[k]);
            } else {
#pragma unroll
                for (uint32_t k = 0; k < 8u; ++k) ST::pair(levels, slot0 + (k & 3u) * rowS + (k >> 2) * zS, own[k], nbv[k]);
            }
        }
        __syncthreads();  // the next half reuses stage
    }
    if (nlev <= 1) return;
    // level 1 (l1 complete): the last column's neighbour is in the next region
    for (uint32_t i = tid; i < 512u; i += 256u) {
        const uint32_t x1 = i & 7u, y1 = (i >> 3) & 7u, z1 = i >> 6;
        const uint64_t idx = brick_addr(vd.lv[1], 8u * rx + x1, 8u * ry + y1, 8u * rz + z1);
        if (x1 < 7u) store_pair<FMT>(levels, idx, l1[i], l1[i + 1u]);
        else if (last_x) store_pair<FMT>(levels, idx, l1[i], l1[i]);
        else store_half<FMT>(levels, idx, 0, l1[i]);
    }
    if (rx > 0u && tid < 64u) {
        const uint32_t y1 = tid & 7u, z1 = tid >> 3;
        store_half<FMT>(levels, brick_addr(vd.lv[1], 8u * rx - 1u, 8u * ry + y1, 8u * rz + z1), 1,
                        l1[(z1 * 8u + y1) * 8u]);
    }
    if (nlev <= 2) return;
    if (tid < 64u) {
        uint32_t x, y, z;
        demorton6(tid, &x, &y, &z);
        float2 s = make_float2(0.0f, 0.0f);
        for (uint32_t dz = 0; dz < 2; ++dz)
            for (uint32_t dy = 0; dy < 2; ++dy)
                for (uint32_t dx = 0; dx < 2; ++dx) s = add2(s, l1[((2u * z + dz) * 8u + 2u * y + dy) * 8u + 2u * x + dx]);
        l2[(z * 4u + y) * 4u + x] = mip_parent2<FMT>(s, 8.0f);
    }
    __syncthreads();
    if (tid < 64u) {
        uint32_t x, y, z;
        demorton6(tid, &x, &y, &z);
        const uint32_t i = (z * 4u + y) * 4u + x;
        const uint64_t idx = brick_addr(vd.lv[2], 4u * rx + x, 4u * ry + y, 4u * rz + z);
        if (x < 3u) store_pair<FMT>(levels, idx, l2[i], l2[i + 1u]);
        else if (last_x) store_pair<FMT>(levels, idx, l2[i], l2[i]);
        else store_half<FMT>(levels, idx, 0, l2[i]);
        if (rx > 0u && x == 0u)
            store_half<FMT>(levels, brick_addr(vd.lv[2], 4u * rx - 1u, 4u * ry + y, 4u * rz + z), 1, l2[i]);
    }
    if (nlev <= 3) return;
    if (tid < 8u) {
        const uint32_t x = tid & 1u, y = (tid >> 1) & 1u, z = tid >> 2;
        float2 s = make_float2(0.0f, 0.0f);
        for (uint32_t dz = 0; dz < 2; ++dz)
            for (uint32_t dy = 0; dy < 2; ++dy)
                for (uint32_t dx = 0; dx < 2; ++dx) s = add2(s, l2[((2u * z + dz) * 4u + 2u * y + dy) * 4u + 2u * x + dx]);
        l3[(z * 2u + y) * 2u + x] = mip_parent2<FMT>(s, 8.0f);
    }
    __syncthreads();
    if (tid < 8u) {
        const uint32_t x = tid & 1u, y = (tid >> 1) & 1u, z = tid >> 2;
        const uint32_t i = (z * 2u + y) * 2u + x;
        const uint64_t idx = brick_addr(vd.lv[3], 2u * rx + x, 2u * ry + y, 2u * rz + z);
        if (x == 0u) store_pair<FMT>(levels, idx, l3[i], l3[i + 1u]);
        else if (last_x) store_pair<FMT>(levels, idx, l3[i], l3[i]);
        else store_half<FMT>(levels, idx, 0, l3[i]);
        if (rx > 0u && x == 0u)
            store_half<FMT>(levels, brick_addr(vd.lv[3], 2u * rx - 1u, 2u * ry + y, 2u * rz + z), 1, l3[i]);
    }
    if (nlev <= 4) return;
    if (tid == 0u) {
        float2 s = make_float2(0.0f, 0.0f);
        for (int c = 0; c < 8; ++c) s = add2(s, l3[c]);
        const float2 p = mip_parent2<FMT>(s, 8.0f);
        const uint64_t idx = brick_addr(vd.lv[4], rx, ry, rz);
        if (last_x) store_pair<FMT>(levels, idx, p, p);
        else store_half<FMT>(levels, idx, 0, p);
        if (rx > 0u) store_half<FMT>(levels, brick_addr(vd.lv[4], rx - 1u, ry, rz), 1, p);
    }
}
// TEX: the density noise comes from vd.ntex (non-null when the volume has
// noise), else it is evaluated here
template <int FMT, bool TEX>
__global__ void __launch_bounds__(256) k_build_fused(GridDesc g, VolumeDesc vd,
                                                     const unsigned long long* __restrict__ shellbits,
                                                     const unsigned long long* __restrict__ fillbits,
                                                     void* __restrict__ levels, uint8_t* __restrict__ zflags, int nlev,
                                                     uint32_t rpa) {
    k_build_fused_body<FMT, TEX>(g, vd, shellbits, fillbits, levels, zflags, nlev, rpa, (int)blockIdx.x,
                                 (int)gridDim.x);
}

// The density noise texture: fbm_block8 of every 2x2x2 block of level-0
// texels (one thread per block, the 8 Morton-consecutive slots of a brick),
// stored in level-0 slot order (brick * 64 + Morton index, from 0); RGBA16F
// volumes keep the binary16 rounding,
// which is all their level 0 uses (a = 1 * noise, rounded to half).
template <int FMT>
__global__ void __launch_bounds__(256) k_noise_tex(GridDesc g, VolumeDesc vd, void* __restrict__ out,
                                                   uint64_t nblocks) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= nblocks) return;
    const uint32_t brick = (uint32_t)(i >> 3), q = (uint32_t)(i & 7u);
    const uint32_t nb = g.nb0;
    const uint32_t bx = brick % nb, by = (brick / nb) % nb, bz = brick / (nb * nb);
    const int x0 = (int)(bx * 4u + 2u * (q & 1u)), y0 = (int)(by * 4u + 2u * ((q >> 1) & 1u));
    const int z0 = (int)(bz * 4u + 2u * (q >> 2));
    float nz[8];
    fbm_block8(vd, g.N, x0, y0, z0, nz);
    if (FMT == 2) {
        uint32_t hw[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) hw[k] = f2h(nz[2 * k]) | (f2h(nz[2 * k + 1]) << 16);
        ((uint4*)out)[i] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
    } else {
        ((float4*)out)[2 * i] = make_float4(nz[0], nz[1], nz[2], nz[3]);
        ((float4*)out)[2 * i + 1] = make_float4(nz[4], nz[5], nz[6], nz[7]);
    }
}

size_t noise_tex_bytes(const GridDesc& g, const VolumeDesc& vd) {
    if (vd.mode != 1 || vd.noise_octaves <= 0 || (vd.format != 1 && vd.format != 2) || g.N % 16 != 0) return 0;
    const uint64_t nb = g.nb0;
    return (size_t)(nb * nb * nb * 64u) * (vd.format == 1 ? 4u : 2u);
}

void launch_noise_tex(const GridDesc& g, const VolumeDesc& vd, void* ntex, hipStream_t s) {
    const uint64_t nb = g.nb0;
    const uint64_t nblocks = nb * nb * nb * 8u;
    const dim3 grid((unsigned)((nblocks + 255) / 256));
    if (vd.format == 2) hipLaunchKernelGGL(k_noise_tex<2>, grid, dim3(256), 0, s, g, vd, ntex, nblocks);
    else hipLaunchKernelGGL(k_noise_tex<1>, grid, dim3(256), 0, s, g, vd, ntex, nblocks);
}

// R8 variant of the fused build (one byte per texel, no neighbour halves)
__global__ void __launch_bounds__(256) k_build_fused_r8(GridDesc g, VolumeDesc vd,
                                                        const unsigned long long* __restrict__ shellbits,
                                                        void* __restrict__ levels, int nlev, uint32_t rpa) {
    __shared__ uint32_t l1[512];
    __shared__ uint32_t l2[64];
    __shared__ uint32_t l3[8];
    const uint32_t r = blockIdx.x;
    const uint32_t rx = r % rpa, ry = (r / rpa) % rpa, rz = r / (rpa * rpa);
    const uint32_t nb0 = g.nb0;
    uint8_t* out = (uint8_t*)levels;
    for (uint32_t rep = 0; rep < 2; ++rep) {
        const uint32_t grp = threadIdx.x + rep * 256u;
        const uint32_t lb = grp >> 3, q = grp & 7u;
        const uint32_t lbx = lb & 3u, lby = (lb >> 2) & 3u, lbz = lb >> 4;
        const uint32_t brick = ((4u * rz + lbz) * nb0 + 4u * ry + lby) * nb0 + 4u * rx + lbx;
        const unsigned long long sh = shellbits[brick];
        uint32_t acc = 0;
#pragma unroll
        for (uint32_t c = 0; c < 8; ++c) {
            const uint32_t m = q * 8u + c;
            const uint32_t v = ((sh >> m) & 1ull) ? 255u : 0u;
            out[vd.lv[0].base + (uint64_t)brick * 64u + m] = (uint8_t)v;
            acc += v;
        }
        if (nlev > 1) {
            const uint32_t p = (acc + 4u) / 8u;
            const uint32_t x1 = 2u * lbx + (q & 1u), y1 = 2u * lby + ((q >> 1) & 1u), z1 = 2u * lbz + (q >> 2);
            out[brick_addr(vd.lv[1], 8u * rx + x1, 8u * ry + y1, 8u * rz + z1)] = (uint8_t)p;
            l1[(z1 * 8u + y1) * 8u + x1] = p;
        }
    }
    if (nlev <= 2) return;
    __syncthreads();
    const uint32_t t = threadIdx.x;
    if (t < 64u) {
        uint32_t x, y, z;
        demorton6(t, &x, &y, &z);
        uint32_t s = 0;
        for (uint32_t dz = 0; dz < 2; ++dz)
            for (uint32_t dy = 0; dy < 2; ++dy)
                for (uint32_t dx = 0; dx < 2; ++dx) s += l1[((2u * z + dz) * 8u + 2u * y + dy) * 8u + 2u * x + dx];
        const uint32_t p = (s + 4u) / 8u;
        out[brick_addr(vd.lv[2], 4u * rx + x, 4u * ry + y, 4u * rz + z)] = (uint8_t)p;
        l2[(z * 4u + y) * 4u + x] = p;
    }
    if (nlev <= 3) return;
    __syncthreads();
    if (t < 8u) {
        const uint32_t x = t & 1u, y = (t >> 1) & 1u, z = t >> 2;
        uint32_t s = 0;
        for (uint32_t dz = 0; dz < 2; ++dz)
            for (uint32_t dy = 0; dy < 2; ++dy)
                for (uint32_t dx = 0; dx < 2; ++dx) s += l2[((2u * z + dz) * 4u + 2u * y + dy) * 4u + 2u * x + dx];
        const uint32_t p = (s + 4u) / 8u;
        out[brick_addr(vd.lv[3], 2u * rx + x, 2u * ry + y, 2u * rz + z)] = (uint8_t)p;
        l3[(z * 2u + y) * 2u + x] = p;
    }
    if (nlev <= 4) return;
    __syncthreads();
    if (t == 0u) {
        uint32_t s = 0;
        for (int c = 0; c < 8; ++c) s += l3[c];
        out[brick_addr(vd.lv[4], rx, ry, rz)] = (uint8_t)((s + 4u) / 8u);
    }
}

// ---------------------------------------------------------------------------
// mip level k from level k-1 (glGenerateMipmap, src/Shaders/VoxelizeShader.cpp:105)
// parent = sum of children (z, y, x order) / count; R8: integer round half up
// ---------------------------------------------------------------------------
__device__ __forceinline__ void child_range(int p, int nparent, int nchild, int* c0, int* c1) {
    *c0 = min(2 * p, nchild - 1);
    *c1 = min(2 * p + 1, nchild - 1);
    if (p == nparent - 1) *c1 = nchild - 1;
}

// (r, a) of RGBA parent (x, y, z) from the children's (r, a)
__device__ __forceinline__ float2 mip_parent_at(const VolumeDesc& vd, const LevelDesc& Cl, const void* levels, int x,
                                                int y, int z, int np) {
    const int nc = (int)Cl.n;
    int x0, x1, y0, y1, z0, z1;
    child_range(x, np, nc, &x0, &x1);
    child_range(y, np, nc, &y0, &y1);
    child_range(z, np, nc, &z0, &z1);
    const int cnt = (x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1);
    float2 s = make_float2(0.0f, 0.0f);
    for (int cz = z0; cz <= z1; ++cz)
        for (int cy = y0; cy <= y1; ++cy)
            for (int cx = x0; cx <= x1; ++cx) s = add2(s, load_ra(levels, vd.format, brick_addr(Cl, cx, cy, cz)));
    return vd.format == 1 ? mip_parent2<1>(s, (float)cnt) : mip_parent2<2>(s, (float)cnt);
}

__global__ void __launch_bounds__(256) k_mip(VolumeDesc vd, int k, void* __restrict__ levels, uint32_t ntexels) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntexels) return;
    const LevelDesc P = vd.lv[k], Cl = vd.lv[k - 1];
    uint32_t brick = t >> 6, m = t & 63u;
    uint32_t bx = brick % P.nb, by = (brick / P.nb) % P.nb, bz = brick / (P.nb * P.nb);
    uint32_t lx, ly, lz;
    demorton6(m, &lx, &ly, &lz);
    int x = (int)(bx * 4u + lx), y = (int)(by * 4u + ly), z = (int)(bz * 4u + lz);
    uint64_t out = P.base + t;
    int np = (int)P.n, nc = (int)Cl.n;
    if (x >= np || y >= np || z >= np) {
        if (vd.format == 0) ((uint8_t*)levels)[out] = 0;
        else store_pair_rt(levels, vd.format, out, make_float2(0.0f, 0.0f), make_float2(0.0f, 0.0f));
        return;
    }
    if (vd.format == 0) {
        int x0, x1, y0, y1, z0, z1;
        child_range(x, np, nc, &x0, &x1);
        child_range(y, np, nc, &y0, &y1);
        child_range(z, np, nc, &z0, &z1);
        int cnt = (x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1);
        uint32_t sum = 0;
        for (int cz = z0; cz <= z1; ++cz)
            for (int cy = y0; cy <= y1; ++cy)
                for (int cx = x0; cx <= x1; ++cx) sum += ((const uint8_t*)levels)[brick_addr(Cl, cx, cy, cz)];
        ((uint8_t*)levels)[out] = (uint8_t)((sum + (uint32_t)cnt / 2u) / (uint32_t)cnt);
        return;
    }
    // the parent and its +x neighbour, each averaged from its own children
    const float2 own = mip_parent_at(vd, Cl, levels, x, y, z, np);
    const float2 nb = x + 1 < np ? mip_parent_at(vd, Cl, levels, x + 1, y, z, np) : own;
    store_pair_rt(levels, vd.format, out, own, nb);
}

// The small end of the chain in one workgroup: levels k0..L-1 (n <= 16, at
// most 4096 texels each), every level kept in LDS for the next one (level
// k0 - 1 is read from memory).  Same sums, order and rounding as k_mip.
constexpr int kTailTexels = 4096;
__device__ __forceinline__ void k_mip_tail_body(VolumeDesc vd, int k0, void* __restrict__ levels) {
    __shared__ float2 buf[2][kTailTexels];  // (r, a), or the R8 value in .x; linear x + n (y + n z)
    const int tid = threadIdx.x;
    int cur = 0;
    for (int k = k0; k < vd.L; ++k) {
        const LevelDesc P = vd.lv[k], Cl = vd.lv[k - 1];
        const int np = (int)P.n, nc = (int)Cl.n;
        const int ntex = np * np * np;
        for (int t = tid; t < ntex; t += blockDim.x) {
            const int x = t % np, y = (t / np) % np, z = t / (np * np);
            int x0, x1, y0, y1, z0, z1;
            child_range(x, np, nc, &x0, &x1);
            child_range(y, np, nc, &y0, &y1);
            child_range(z, np, nc, &z0, &z1);
            const int cnt = (x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1);
            if (vd.format == 0) {
                uint32_t sum = 0;
                for (int cz = z0; cz <= z1; ++cz)
                    for (int cy = y0; cy <= y1; ++cy)
                        for (int cx = x0; cx <= x1; ++cx)
                            sum += k == k0 ? (uint32_t)((const uint8_t*)levels)[brick_addr(Cl, cx, cy, cz)]
                                           : (uint32_t)buf[cur ^ 1][cx + nc * (cy + nc * cz)].x;
                buf[cur][t] = make_float2((float)((sum + (uint32_t)cnt / 2u) / (uint32_t)cnt), 0.0f);
            } else {
                float2 sm = make_float2(0.0f, 0.0f);
                for (int cz = z0; cz <= z1; ++cz)
                    for (int cy = y0; cy <= y1; ++cy)
                        for (int cx = x0; cx <= x1; ++cx)
                            sm = add2(sm, k == k0 ? load_ra(levels, vd.format, brick_addr(Cl, cx, cy, cz))
                                                  : buf[cur ^ 1][cx + nc * (cy + nc * cz)]);
                buf[cur][t] = vd.format == 1 ? mip_parent2<1>(sm, (float)cnt) : mip_parent2<2>(sm, (float)cnt);
            }
        }
        __syncthreads();
        // every slot of the level's bricks (zeros beyond the side)
        const uint32_t nslots = P.nb * P.nb * P.nb * 64u;
        for (uint32_t t = tid; t < nslots; t += blockDim.x) {
            const uint32_t brick = t >> 6, m = t & 63u;
            uint32_t lx, ly, lz;
            demorton6(m, &lx, &ly, &lz);
            const int x = (int)((brick % P.nb) * 4u + lx), y = (int)(((brick / P.nb) % P.nb) * 4u + ly);
            const int z = (int)((brick / (P.nb * P.nb)) * 4u + lz);
            const uint64_t out = P.base + t;
            const bool in = x < np && y < np && z < np;
            const int li = x + np * (y + np * z);
            if (vd.format == 0) {
                ((uint8_t*)levels)[out] = in ? (uint8_t)buf[cur][li].x : (uint8_t)0;
            } else {
                const float2 own = in ? buf[cur][li] : make_float2(0.0f, 0.0f);
                const float2 nb = in && x + 1 < np ? buf[cur][li + 1] : own;
                store_pair_rt(levels, vd.format, out, own, nb);
            }
        }
        cur ^= 1;
        __syncthreads();
    }
}
__global__ void __launch_bounds__(1024) k_mip_tail(VolumeDesc vd, int k0, void* __restrict__ levels) {
    k_mip_tail_body(vd, k0, levels);
}

// ---------------------------------------------------------------------------
// camera pass setup: world/view centres, draw-order key, tile box
// ---------------------------------------------------------------------------
// sort key of billboard j for CloudVolume::sortBoards(point): the distance
// glm::distance(position + p_j, point) (src/CloudVolume.cpp:69) with its bits
// inverted above the index, so an ascending sort of keys is far to near with
// the lower index first among equal distances
__device__ __forceinline__ unsigned long long board_key(float px, float py, float pz, float cx, float cy, float cz,
                                                        int j) {
    float ex = px - cx, ey = py - cy, ez = pz - cz;
    float dist = sqrtf((ex * ex + ey * ey) + ez * ez);
    return ((unsigned long long)(~__float_as_uint(dist)) << 32) | (unsigned long long)(uint32_t)j;
}

__global__ void __launch_bounds__(256) k_sort_keys(GridDesc g, const float4* __restrict__ boards, int S, float px,
                                                   float py, float pz, unsigned long long* __restrict__ keys) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= S) return;
    float4 b = boards[j];
    keys[j] = board_key(px, py, pz, g.vpos[0] + b.x, g.vpos[1] + b.y, g.vpos[2] + b.z, j);
}

__global__ void __launch_bounds__(256) k_iota(int32_t* __restrict__ order, int S) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < S) order[j] = j;
}

__device__ __forceinline__ void k_cam_setup_body(GridDesc g, CameraView cv, const float4* __restrict__ boards, int S,
                                                   float4* __restrict__ sph_world, float4* __restrict__ sph_view,
                                                   int4* __restrict__ sph_tiles, unsigned long long* __restrict__ keys, int bid, int nblk) {
    int j = bid * blockDim.x + threadIdx.x;
    if (j >= S) return;
    float4 b = boards[j];
    float cx = g.vpos[0] + b.x, cy = g.vpos[1] + b.y, cz = g.vpos[2] + b.z, s = b.w;
    float vx = ((cv.V[0] * cx + cv.V[4] * cy) + cv.V[8] * cz) + cv.V[12];
    float vy = ((cv.V[1] * cx + cv.V[5] * cy) + cv.V[9] * cz) + cv.V[13];
    float vz = ((cv.V[2] * cx + cv.V[6] * cy) + cv.V[10] * cz) + cv.V[14];
    sph_world[j] = make_float4(cx, cy, cz, s);
    sph_view[j] = make_float4(vx, vy, vz, 0.0f);
    // glm::distance(position + p, camera) (src/CloudVolume.cpp:69)
    keys[j] = board_key(cv.eye[0], cv.eye[1], cv.eye[2], cx, cy, cz, j);
    float dz = -vz;
    int4 tiles = make_int4(1, 0, 1, 0);  // empty
    if ((s > 0.0f) && (dz >= cv.znear && dz <= cv.zfar)) {
        float ex_ = cv.kx * dz, ey_ = cv.ky * dz;
        float xa = ((vx - s) / ex_ + 1.0f) * 0.5f * (float)cv.W - 0.5f;
        float xb = ((vx + s) / ex_ + 1.0f) * 0.5f * (float)cv.W - 0.5f;
        float ya = ((vy - s) / ey_ + 1.0f) * 0.5f * (float)cv.H - 0.5f;
        float yb = ((vy + s) / ey_ + 1.0f) * 0.5f * (float)cv.H - 0.5f;
        float fx0 = floorf(fminf(xa, xb)) - 1.0f, fx1 = ceilf(fmaxf(xa, xb)) + 1.0f;
        float fy0 = floorf(fminf(ya, yb)) - 1.0f, fy1 = ceilf(fmaxf(ya, yb)) + 1.0f;
        if (fx1 >= 0.0f && fy1 >= 0.0f && fx0 <= (float)(cv.W - 1) && fy0 <= (float)(cv.H - 1)) {
            int x0 = (int)fmaxf(fx0, 0.0f), x1 = (int)fminf(fx1, (float)(cv.W - 1));
            int y0 = (int)fmaxf(fy0, 0.0f), y1 = (int)fminf(fy1, (float)(cv.H - 1));
            tiles = make_int4(x0 / cv.tile_w, x1 / cv.tile_w, y0 / cv.tile_h, y1 / cv.tile_h);
        }
    }
    sph_tiles[j] = tiles;
}
__global__ void __launch_bounds__(256) k_cam_setup(GridDesc g, CameraView cv, const float4* __restrict__ boards, int S,
                                                   float4* __restrict__ sph_world, float4* __restrict__ sph_view,
                                                   int4* __restrict__ sph_tiles, unsigned long long* __restrict__ keys) {
    k_cam_setup_body(g, cv, boards, S, sph_world, sph_view, sph_tiles, keys, (int)blockIdx.x, (int)gridDim.x);
}

// ---------------------------------------------------------------------------
// draw order (CloudVolume::sortBoards, src/CloudVolume.cpp:65-82)
//
// The reference runs a selection sort: for i, m = first position of the
// maximum distance in [i, S), swap(i, m).  Keys (~distance bits << 32 | index)
// sorted ascending give the distinct distances far to near.  Equal distances
// are frequent in FP32 (thousands of billboards at similar range), and for
// them the reference's order depends on its swap history: inside a run of
// equal distances it takes the member with the lowest *current* position.
// replay_ties() reproduces that exactly in O(S): positions are tracked
// through every swap (pos: element -> position, at: position -> element)
// and only the runs of equal distances scan their members.
// ---------------------------------------------------------------------------
template <class I, class P>
__device__ void replay_ties(int S, const I* s_idx, const uint32_t* s_tie, P* pos, P* at,
                            int32_t* __restrict__ order) {
    auto place = [&](int k, int sel) {
        int m = (int)pos[sel];
        int e = (int)at[k];
        at[k] = (P)sel;
        at[m] = (P)e;
        pos[sel] = (P)k;
        pos[e] = (P)m;
        order[k] = sel;
    };
    int i = 0;
    while (i < S) {
        int j = i + 1;
        while (j < S && ((s_tie[j >> 5] >> (j & 31)) & 1u)) ++j;
        if (j == i + 1) {
            place(i, (int)s_idx[i]);
            ++i;
            continue;
        }
        for (int k = i; k < j; ++k) {
            int best = -1, bp = 0x7fffffff;
            for (int t = i; t < j; ++t) {
                int e = (int)s_idx[t];
                int p = (int)pos[e];
                if (p >= k && p < bp) { bp = p; best = e; }
            }
            place(k, best);
        }
        i = j;
    }
}

// Parallel form of the same replay for one workgroup.  Let D[k] be the
// position the element selected at step k held just before its swap.  An
// unplaced element y sits at position y until step y; whenever the scan
// reaches its position p it is displaced to D[p].  Hence its position at step
// k is find(y, k): x = y; while (x < k) x = D[x].  Steps outside runs of equal
// distances select sorted[k] whatever the positions, so D[k] = find(sorted[k], k)
// for a whole segment of such steps is computed in parallel rounds (a chain
// only waits on smaller steps of the same segment).  A run of equal distances
// is resolved step by step: the unplaced member with the smallest find() wins.
// s_idx[t]: bits 0-14 element, bit 15 = same distance as entry t-1;
// s_D: positions, 0xFFFF = not yet known; placed: one bit per element.
__device__ void replay_ties_block(int S, const uint16_t* s_idx, uint16_t* s_D, uint32_t* placed,
                                  int32_t* __restrict__ order) {
    __shared__ int seg_end;
    __shared__ int red_p[16], red_e[16];
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wid = tid >> 6;
    auto tie_at = [&](int t) -> bool { return (s_idx[t] >> 15) != 0; };
    auto elem = [&](int t) -> int { return (int)(s_idx[t] & 0x7fffu); };
    int k0 = 0;
    while (k0 < S) {
        if (k0 + 1 < S && tie_at(k0 + 1)) {
            // run of equal distances [k0, ge)
            int ge = k0 + 1;
            while (ge < S && tie_at(ge)) ++ge;
            for (int k = k0; k < ge; ++k) {
                int best_p = 0x7fffffff, best_e = -1;
                for (int t = k0 + tid; t < ge; t += nt) {
                    int y = elem(t);
                    if ((placed[y >> 5] >> (y & 31)) & 1u) continue;
                    int x = y;
                    while (x < k) x = (int)s_D[x];
                    if (x < best_p) { best_p = x; best_e = y; }
                }
                for (int off = 32; off > 0; off >>= 1) {
                    int op = __shfl_xor(best_p, off, 64);
                    int oe = __shfl_xor(best_e, off, 64);
                    if (op < best_p) { best_p = op; best_e = oe; }
                }
                if (lane == 0) { red_p[wid] = best_p; red_e[wid] = best_e; }
                __syncthreads();
                if (tid == 0) {
                    int p = red_p[0], e = red_e[0];
                    for (int w = 1; w < nt / 64; ++w)
                        if (red_p[w] < p) { p = red_p[w]; e = red_e[w]; }
                    if (e >= 0) {  // always: ge - k members are unplaced
                        s_D[k] = (uint16_t)p;
                        order[k] = e;
                        placed[e >> 5] |= 1u << (e & 31);
                    }
                }
                __syncthreads();
            }
            k0 = ge;
            continue;
        }
        // segment [k0, kend): steps outside runs of equal distances
        if (tid == 0) seg_end = min(k0 + nt, S);
        __syncthreads();
        {
            int k = k0 + tid;
            if (k < min(k0 + nt, S) && (tie_at(k) || (k + 1 < S && tie_at(k + 1)))) atomicMin(&seg_end, k);
        }
        __syncthreads();
        const int kend = seg_end;
        const int k = k0 + tid;
        bool done = k >= kend;
        int x = done ? 0 : elem(k);
        while (true) {
            if (!done) {
                while (true) {
                    while (x < k0) x = (int)s_D[x];
                    if (x >= k) {
                        s_D[k] = (uint16_t)x;
                        order[k] = elem(k);
                        done = true;
                        break;
                    }
                    int dx = (int)s_D[x];  // x in [k0, k): known once its own chain finished
                    if (dx == 0xFFFF) break;
                    x = dx;
                }
            }
            if (__syncthreads_count(done ? 0 : 1) == 0) break;
        }
        k0 = kend;
    }
}

__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int m) {
    const int lo = __shfl_xor((int)(uint32_t)v, m, 64), hi = __shfl_xor((int)(uint32_t)(v >> 32), m, 64);
    return ((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Bitonic sort of the wave's 64 * E keys held in registers, key r = lane * E + e
// (ascending): partners within a lane swap registers, partners in other
// lanes are exchanged with cross-lane shuffles; no LDS and no barriers.
template <int E>
__device__ __forceinline__ void wave_bitonic(unsigned long long (&v)[E], int lane) {
    constexpr int n = 64 * E;
#pragma unroll
    for (int k = 2; k <= n; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j < E) {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int p = e ^ j;
                    if (p > e) {
                        const bool up = ((lane * E + e) & k) == 0;
                        const unsigned long long a = v[e], b = v[p];
                        const bool sw = up ? (a > b) : (a < b);
                        v[e] = sw ? b : a;
                        v[p] = sw ? a : b;
                    }
                }
            } else {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int r = lane * E + e;
                    const bool up = (r & k) == 0;
                    const unsigned long long o = shfl_xor_u64(v[e], j / E);
                    const bool keep_min = ((r & j) == 0) == up;
                    const unsigned long long mn = o < v[e] ? o : v[e], mx = o < v[e] ? v[e] : o;
                    v[e] = keep_min ? mn : mx;
                }
            }
        }
    }
}

// Sort of npad = 1024 * E keys (the first S real, the rest ~0) in one
// workgroup of 1024: each wave bitonic-sorts its 64 * E keys in registers,
// then four merge rounds pair the 16 sorted runs through LDS (merge path:
// a binary search gives each thread its E outputs).  The keys are
// (~distance bits << 32 | index), distinct for distinct billboards, so the
// ascending order is the draw order up to runs of equal distances.
template <int E>
__device__ __forceinline__ void k_sort_lds_body(const unsigned long long* __restrict__ keys_in, int S,
                                                   int32_t* __restrict__ order, int* __restrict__ tie_flag) {
    extern __shared__ unsigned long long sk[];
    __shared__ int tie;
    constexpr int npad = 1024 * E;
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63;
    if (tid == 0) tie = 0;
    {
        unsigned long long v[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int i = tid * E + e;
            v[e] = i < S ? keys_in[i] : ~0ull;
        }
        wave_bitonic<E>(v, lane);
#pragma unroll
        for (int e = 0; e < E; ++e) sk[tid * E + e] = v[e];
    }
    __syncthreads();
#pragma unroll 1
    for (int L = 64 * E; L < npad; L <<= 1) {
        const int pos = tid * E;
        const int base = pos & ~(2 * L - 1);
        const int o = pos - base;
        const unsigned long long* A = sk + base;
        const unsigned long long* B = A + L;
        // i = how many of the first o merged keys come from A (ties: A first)
        int lo = max(0, o - L), hi = min(o, L);
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (A[mid] <= B[o - 1 - mid]) lo = mid + 1;
            else hi = mid;
        }
        int ia = lo, ib = o - lo;
        unsigned long long out[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const unsigned long long a = ia < L ? A[ia] : ~0ull, b = ib < L ? B[ib] : ~0ull;
            const bool take_a = ib >= L || (ia < L && a <= b);
            out[e] = take_a ? a : b;
            ia += take_a ? 1 : 0;
            ib += take_a ? 0 : 1;
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) sk[pos + e] = out[e];
        __syncthreads();
    }
    // Reuse the key storage: idx u16[npad] (bit 15 = same distance as the
    // previous key) | pos u16[npad] | at u16[npad].  Chunk c of 1024 keys is
    // read before its indices are written; those writes cover keys
    // [256 c, 256 c + 256), all of which were read in chunks <= c, and the
    // previous key of the next chunk (index 1024 (c+1) - 1) is not touched.
    uint16_t* s_idx = (uint16_t*)sk;
    uint16_t* s_pos = s_idx + npad;
    uint16_t* s_at = s_pos + npad;
    for (int c = 0; c < npad; c += nt) {
        const int i = c + tid;
        uint32_t v = 0;
        if (i < S) {
            unsigned long long k = sk[i];
            v = (uint32_t)(k & 0xffffffffull);
            if (i > 0 && (k >> 32) == (sk[i - 1] >> 32)) {
                v |= 0x8000u;
                tie = 1;
            }
        }
        __syncthreads();
        if (i < S) s_idx[i] = (uint16_t)v;
        __syncthreads();
    }
    if (!tie) {
        for (int i = tid; i < S; i += nt) order[i] = (int32_t)s_idx[i];
        if (tid == 0) *tie_flag = 0;
        return;
    }
    __shared__ uint32_t placed[kLdsSortMax / 32];
    uint16_t* s_D = s_pos;
    (void)s_at;
    for (int i = tid; i < S; i += nt) s_D[i] = 0xFFFFu;
    for (int w = tid; w < kLdsSortMax / 32; w += nt) placed[w] = 0u;
    __syncthreads();
    replay_ties_block(S, s_idx, s_D, placed, order);
    if (tid == 0) *tie_flag = 1;
}
template <int E>
__global__ void __launch_bounds__(1024) k_sort_lds(const unsigned long long* __restrict__ keys_in, int S,
                                                   int32_t* __restrict__ order, int* __restrict__ tie_flag) {
    k_sort_lds_body<E>(keys_in, S, order, tie_flag);
}

// global-memory bitonic stage for more billboards than one LDS holds
__global__ void k_bitonic_global(unsigned long long* __restrict__ keys, int npad, int j, int k) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    int ixj = i ^ j;
    if (ixj > i) {
        unsigned long long a = keys[i], b = keys[ixj];
        bool up = (i & k) == 0;
        if ((a > b) == up) { keys[i] = b; keys[ixj] = a; }
    }
}
__global__ void k_pad_keys(const unsigned long long* __restrict__ in, unsigned long long* __restrict__ out, int S,
                           int npad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npad) out[i] = i < S ? in[i] : ~0ull;
}
// tie bits / u16 indices of the global path, then the replay in global memory
__global__ void k_sort_finish(const unsigned long long* __restrict__ sk, int S, int32_t* __restrict__ order,
                              uint32_t* __restrict__ g_idx, uint32_t* __restrict__ g_tie, int* __restrict__ tie_flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    uint32_t e = (uint32_t)(sk[i] & 0xffffffffull);
    order[i] = (int32_t)e;
    g_idx[i] = e;
    if (i > 0 && (sk[i] >> 32) == (sk[i - 1] >> 32)) {
        atomicOr(&g_tie[i >> 5], 1u << (i & 31));
        atomicOr(tie_flag, 1);
    }
}
__global__ void k_sort_replay_global(int S, const uint32_t* __restrict__ g_idx, const uint32_t* __restrict__ g_tie,
                                     int32_t* __restrict__ pos, int32_t* __restrict__ at, int32_t* __restrict__ order,
                                     const int* __restrict__ tie_flag) {
    if (*tie_flag == 0) return;
    for (int i = 0; i < S; ++i) { pos[i] = i; at[i] = i; }
    replay_ties<uint32_t, int32_t>(S, g_idx, g_tie, pos, at, order);
}

// ---------------------------------------------------------------------------
// binning: set bit `rank` in every tile a billboard's box overlaps; one
// wave per billboard, lanes striding over its tile box (a near billboard
// covers ~1000 tiles, which one thread would walk alone)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void k_bin_body(const int4* __restrict__ sph_tiles, const int32_t* __restrict__ order, int S,
                                             int tiles_x, int words, unsigned long long* __restrict__ masks,
                                             int* __restrict__ tile_count, int bid, int nblk) {
    const int r = bid * 4 + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= S) return;
    const int4 t = sph_tiles[order[r]];
    if (t.x > t.y || t.z > t.w) return;
    const int nx = t.y - t.x + 1;
    const int n = nx * (t.w - t.z + 1);
    const unsigned long long bit = 1ull << (r & 63);
    const int w = r >> 6;
    for (int k = lane; k < n; k += 64) {
        const int ty = t.z + k / nx, tx = t.x + k % nx;
        atomicOr(&masks[((size_t)ty * tiles_x + tx) * words + w], bit);
        atomicAdd(&tile_count[ty * tiles_x + tx], 1);
    }
}
__global__ void __launch_bounds__(256) k_bin(const int4* __restrict__ sph_tiles, const int32_t* __restrict__ order, int S,
                                             int tiles_x, int words, unsigned long long* __restrict__ masks,
                                             int* __restrict__ tile_count) {
    k_bin_body(sph_tiles, order, S, tiles_x, words, masks, tile_count, (int)blockIdx.x, (int)gridDim.x);
}

// Longest job first: tiles are bucketed by log2(number of binned billboards)
// and handed to the persistent cone waves heaviest bucket first, so the
// expensive tiles of clump centres do not start last and form the tail.
__global__ void __launch_bounds__(256) k_tile_bucket(const int* __restrict__ tile_count, int ntiles,
                                                     int* __restrict__ bucket_of, int* __restrict__ hist) {
    __shared__ int lh[32];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x < 32) lh[threadIdx.x] = 0;
    __syncthreads();
    if (t < ntiles) {
        int c = tile_count[t];
        int key = 31 - (c == 0 ? 0 : 32 - __clz(c));  // heavy tiles get small keys
        bucket_of[t] = key;
        atomicAdd(&lh[key], 1);
    }
    __syncthreads();
    if (threadIdx.x < 32 && lh[threadIdx.x] > 0) atomicAdd(&hist[threadIdx.x], lh[threadIdx.x]);
}
// counting-sort scatter: ranks inside a workgroup from LDS atomics, one global
// atomic per (workgroup, bucket) to reserve the range
__global__ void __launch_bounds__(256) k_tile_scatter(const int* __restrict__ bucket_of, int ntiles,
                                                      const int* __restrict__ hist, int* __restrict__ cursor,
                                                      int* __restrict__ tile_order) {
    __shared__ int lh[32], lbase[32];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x < 32) lh[threadIdx.x] = 0;
    __syncthreads();
    const int key = t < ntiles ? bucket_of[t] : -1;
    const int lrank = key >= 0 ? atomicAdd(&lh[key], 1) : 0;
    __syncthreads();
    if (threadIdx.x < 32 && lh[threadIdx.x] > 0) lbase[threadIdx.x] = atomicAdd(&cursor[threadIdx.x], lh[threadIdx.x]);
    __syncthreads();
    if (key < 0) return;
    int base = 0;
    for (int k = 0; k < key; ++k) base += hist[k];
    tile_order[base + lbase[key] + lrank] = t;
}

// ---------------------------------------------------------------------------
// cone trace (res/conetrace_frag.glsl:64-79, 176-200) + compositing
// ---------------------------------------------------------------------------
// Texel layout of the RGBA formats.  Every RGBA volume is grey (r = g = b:
// the stores are (1,1,1,1), (0,0,0,1) or (s,s,s,s), the mips average the
// channels alike and noise scales only a), so a texel slot holds (r, a) of
// its texel and (r, a) of the +x neighbour x+ = min(x + 1, n - 1):
//   RGBA32F: float4 (r, a, r+, a+);  RGBA16F: halves (r, a, r+, a+).
// A trilinear footprint then takes one slot per y/z corner (4 loads, not 8);
// readback rebuilds RGBA as (r, r, r, a).

// r of one texel (R8, or the r of an RGBA slot)
template <int FMT>
__device__ __forceinline__ float fetch1(const void* __restrict__ vol, uint64_t i) {
    if (FMT == 0) return (float)((const uint8_t*)vol)[i] / 255.0f;
    if (FMT == 1) return ((const float*)vol)[i * 4];
    return h2f(((const unsigned short*)vol)[i * 4]);
}

// Clamped corner coordinates of one axis.  Texel indices inside a level are
// 32-bit (the side is at most 1624: 406^3 bricks of 64 texels < 2^32) and
// the brick-row and brick-slab strides fit the 24-bit multiplier
// (64 * 406^2 < 2^24).
struct AxisTaps {
    float u;      // texel-space coordinate - 0.5
    float fl;     // floor(u)
    int i;        // (int)fl
    uint32_t c0, c1;  // min(max(i, 0), n - 1), min(max(i + 1, 0), n - 1)
};
__device__ __forceinline__ AxisTaps axis_taps(float s, float fn, int nm1) {
    AxisTaps a;
    a.u = fmaf(fminf(fmaxf(s, -1.0f), 2.0f), fn, -0.5f);
    a.fl = floorf(a.u);
    a.i = (int)a.fl;
    a.c0 = (uint32_t)min(max(a.i, 0), nm1);
    a.c1 = (uint32_t)min(max(a.i + 1, 0), nm1);
    return a;
}
// the same for s in [0, 1] (every front-to-back sample): the coordinate
// clamp is the identity and i lies in [-1, n - 1], so one bound per corner.
// w is u - floor(u) from v_fract_f32: exact for u >= 0 (Sterbenz), and for
// u in [-0.5, 0) both corners clamp to texel 0, where any weight returns
// that texel exactly; i is floor(u) converted in one instruction.
struct AxisTapsIn {
    float w;
    int i;
    uint32_t c0, c1;
};
__device__ __forceinline__ AxisTapsIn axis_taps_in(float s, float fn, int nm1) {
    AxisTapsIn a;
    const float u = fmaf(s, fn, -0.5f);
    a.w = __builtin_amdgcn_fractf(u);
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(a.i) : "v"(u));  // (int)floor(u), u finite and small
    a.c0 = (uint32_t)max(a.i, 0);
    a.c1 = (uint32_t)min(a.i + 1, nm1);
    return a;
}

// GL trilinear filtering of r in one level with CLAMP_TO_EDGE, lerps x, y, z
template <int FMT>
__device__ __forceinline__ float trilinear_r(const void* __restrict__ vol, const LevelDesc& L, float sx, float sy,
                                             float sz) {
    const float fn = L.fn;
    const int nm1 = (int)L.n - 1;
    const AxisTaps ax = axis_taps(sx, fn, nm1), ay = axis_taps(sy, fn, nm1), az = axis_taps(sz, fn, nm1);
    const float wx = ax.u - ax.fl, wy = ay.u - ay.fl, wz = az.u - az.fl;
    const uint32_t srow = L.nb << 6, sslab = (L.nb * L.nb) << 6;  // texels per brick row / slab (< 2^24)
    const uint32_t X0 = morton_off_x(ax.c0), X1 = morton_off_x(ax.c1);
    const uint32_t Y0 = morton_off_y(ay.c0, srow), Y1 = morton_off_y(ay.c1, srow);
    const uint32_t Z0 = morton_off_z(az.c0, sslab), Z1 = morton_off_z(az.c1, sslab);
    const uint64_t base = L.base;
    float t000 = fetch1<FMT>(vol, base + (X0 + Y0 + Z0)), t100 = fetch1<FMT>(vol, base + (X1 + Y0 + Z0));
    float t010 = fetch1<FMT>(vol, base + (X0 + Y1 + Z0)), t110 = fetch1<FMT>(vol, base + (X1 + Y1 + Z0));
    float t001 = fetch1<FMT>(vol, base + (X0 + Y0 + Z1)), t101 = fetch1<FMT>(vol, base + (X1 + Y0 + Z1));
    float t011 = fetch1<FMT>(vol, base + (X0 + Y1 + Z1)), t111 = fetch1<FMT>(vol, base + (X1 + Y1 + Z1));
    float c00 = lerpF(t000, t100, wx), c10 = lerpF(t010, t110, wx);
    float c01 = lerpF(t001, t101, wx), c11 = lerpF(t011, t111, wx);
    float c0 = lerpF(c00, c10, wy), c1 = lerpF(c01, c11, wy);
    return lerpF(c0, c1, wz);
}

// Grey sampling split into fetch and filter, so a two-level sample issues
// the loads of both footprints before waiting for any of them.
struct GreyTaps {
    float2 t[8];  // RGBA32F: (r, a) at x0y0z0, x1y0z0, x0y1z0, x1y1z0, x0y0z1, x1y0z1, x0y1z1, x1y1z1
    uint2 h[4];   // RGBA16F: the four slots (x0y0z0, x0y1z0, x0y0z1, x0y1z1) as loaded
    float ax, ay, az;
};
// (float)b - (float)a of the low / high binary16 halves of two words, rounded
// once: 1 * b + (-a) on v_fma_mix_f32 (the conversions are exact)
__device__ __forceinline__ float hdiff_lo(uint32_t b, uint32_t a) {
    float r;
    asm("v_fma_mix_f32 %0, 1.0, %1, -%2 op_sel_hi:[0,1,1]" : "=v"(r) : "v"(b), "v"(a));
    return r;
}
__device__ __forceinline__ float hdiff_hi(uint32_t b, uint32_t a) {
    float r;
    asm("v_fma_mix_f32 %0, 1.0, %1, -%2 op_sel:[0,1,1] op_sel_hi:[0,1,1]" : "=v"(r) : "v"(b), "v"(a));
    return r;
}
// (sx, sy, sz) in [0, 1]^3: callers stop a cone before it leaves the volume
template <int FMT>
__device__ __forceinline__ void grey_fetch(const void* __restrict__ vol, const LevelDesc& L, float sx, float sy,
                                           float sz, GreyTaps& T) {
    static_assert(FMT == 1 || FMT == 2, "RGBA formats");
    const float fn = L.fn;
    const int nm1 = (int)L.n - 1;
    const AxisTapsIn ax = axis_taps_in(sx, fn, nm1), ay = axis_taps_in(sy, fn, nm1), az = axis_taps_in(sz, fn, nm1);
    // x1 is the slot's neighbour x0+ except below the first texel, where both
    // x corners clamp to texel 0: there the weight 0 returns texel 0 exactly,
    // as lerp(t0, t0, w) does
    T.ax = ax.i < 0 ? 0.0f : ax.w;
    T.ay = ay.w;
    T.az = az.w;
    const uint32_t srow = L.nb << 6, sslab = (L.nb * L.nb) << 6;  // texels per brick row / slab (< 2^24)
    const uint32_t X0 = morton_off_x(ax.c0);
    const uint32_t Y0 = morton_off_y(ay.c0, srow), Y1 = morton_off_y(ay.c1, srow);
    const uint32_t Z0 = morton_off_z(az.c0, sslab), Z1 = morton_off_z(az.c1, sslab);
    if constexpr (FMT == 1) {
        const float4* v = (const float4*)vol + L.base;
        const float4 q0 = v[X0 + Y0 + Z0], q1 = v[X0 + Y1 + Z0], q2 = v[X0 + Y0 + Z1], q3 = v[X0 + Y1 + Z1];
        T.t[0] = make_float2(q0.x, q0.y); T.t[1] = make_float2(q0.z, q0.w);
        T.t[2] = make_float2(q1.x, q1.y); T.t[3] = make_float2(q1.z, q1.w);
        T.t[4] = make_float2(q2.x, q2.y); T.t[5] = make_float2(q2.z, q2.w);
        T.t[6] = make_float2(q3.x, q3.y); T.t[7] = make_float2(q3.z, q3.w);
    } else {
        const uint2* v = (const uint2*)vol + L.base;
        T.h[0] = v[X0 + Y0 + Z0]; T.h[1] = v[X0 + Y1 + Z0]; T.h[2] = v[X0 + Y0 + Z1]; T.h[3] = v[X0 + Y1 + Z1];
    }
}
// lerps along x, then y, then z (GL's trilinear order)
__device__ __forceinline__ float2 grey_filter(const GreyTaps& T) {
    float2 o;
    {
        float c00 = lerpF(T.t[0].x, T.t[1].x, T.ax), c10 = lerpF(T.t[2].x, T.t[3].x, T.ax);
        float c01 = lerpF(T.t[4].x, T.t[5].x, T.ax), c11 = lerpF(T.t[6].x, T.t[7].x, T.ax);
        float c0 = lerpF(c00, c10, T.ay), c1 = lerpF(c01, c11, T.ay);
        o.x = lerpF(c0, c1, T.az);
    }
    {
        float c00 = lerpF(T.t[0].y, T.t[1].y, T.ax), c10 = lerpF(T.t[2].y, T.t[3].y, T.ax);
        float c01 = lerpF(T.t[4].y, T.t[5].y, T.ax), c11 = lerpF(T.t[6].y, T.t[7].y, T.ax);
        float c0 = lerpF(c00, c10, T.ay), c1 = lerpF(c01, c11, T.ay);
        o.y = lerpF(c0, c1, T.az);
    }
    return o;
}
// RGBA16F: the x lerps straight from the halves, lerp(a, b, w) =
// fma(w, b - a, a) with b - a and the fma on v_fma_mix_f32 (the same
// roundings as on converted floats); then y and z as grey_filter
__device__ __forceinline__ float2 grey_filter_h(const GreyTaps& T) {
    float cr[4], ca[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t a = T.h[k].x, b = T.h[k].y;  // (r, a) at x0 and at x1
        cr[k] = fmaf(T.ax, hdiff_lo(b, a), h2f(a & 0xffffu));
        ca[k] = fmaf(T.ax, hdiff_hi(b, a), h2f(a >> 16));
    }
    float2 o;
    o.x = lerpF(lerpF(cr[0], cr[1], T.ay), lerpF(cr[2], cr[3], T.ay), T.az);
    o.y = lerpF(lerpF(ca[0], ca[1], T.ay), lerpF(ca[2], ca[3], T.ay), T.az);
    return o;
}
template <int FMT>
__device__ __forceinline__ float2 grey_filter_f(const GreyTaps& T) {
    if constexpr (FMT == 2) return grey_filter_h(T);
    else return grey_filter(T);
}

// Table-addressed grey fetch (the cone kernel with per-level LDS tables):
// the corners and loads of grey_fetch, their byte offsets read from the
// level's tables (ConeTab) instead of computed from Morton offsets.  tx, ty,
// tz point at the entries of floor(u) = 0; lvl is the level's first byte.
template <int FMT>
__device__ __forceinline__ void grey_fetch_tab(const char* __restrict__ lvl, const uint32_t* tx, const uint2* ty,
                                               const uint2* tz, float fn, float sx, float sy, float sz, GreyTaps& T) {
    static_assert(FMT == 1 || FMT == 2, "RGBA formats");
    const float ux = fmaf(sx, fn, -0.5f), uy = fmaf(sy, fn, -0.5f), uz = fmaf(sz, fn, -0.5f);
    int ix, iy, iz;  // floor(u) in [-1, n - 1] for (sx, sy, sz) in [0, 1]^3
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(ix) : "v"(ux));
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(iy) : "v"(uy));
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(iz) : "v"(uz));
    T.ax = ix < 0 ? 0.0f : __builtin_amdgcn_fractf(ux);  // as grey_fetch: texel 0 below the first centre
    T.ay = __builtin_amdgcn_fractf(uy);
    T.az = __builtin_amdgcn_fractf(uz);
    const uint32_t X = tx[ix];
    const uint2 Y = ty[iy], Z = tz[iz];
    const uint32_t o0 = X + Y.x + Z.x, o1 = X + Y.y + Z.x, o2 = X + Y.x + Z.y, o3 = X + Y.y + Z.y;
    if constexpr (FMT == 1) {
        const float4 q0 = *(const float4*)(lvl + o0), q1 = *(const float4*)(lvl + o1);
        const float4 q2 = *(const float4*)(lvl + o2), q3 = *(const float4*)(lvl + o3);
        T.t[0] = make_float2(q0.x, q0.y); T.t[1] = make_float2(q0.z, q0.w);
        T.t[2] = make_float2(q1.x, q1.y); T.t[3] = make_float2(q1.z, q1.w);
        T.t[4] = make_float2(q2.x, q2.y); T.t[5] = make_float2(q2.z, q2.w);
        T.t[6] = make_float2(q3.x, q3.y); T.t[7] = make_float2(q3.z, q3.w);
    } else {
        T.h[0] = *(const uint2*)(lvl + o0); T.h[1] = *(const uint2*)(lvl + o1);
        T.h[2] = *(const uint2*)(lvl + o2); T.h[3] = *(const uint2*)(lvl + o3);
    }
}

// One level's tables and first byte, as the table-addressed march holds them
struct TabLevel {
    const char* lvl;
    const uint32_t* tx;
    const uint2* ty;
    const uint2* tz;
};
// the LDS tables of a workgroup and their per-level offsets (kernel argument)
struct TabRef {
    const uint32_t* lds;
    const ConeTab* t;
};
template <int FMT>
__device__ __forceinline__ TabLevel tab_level(const TabRef& tr, const void* vol, const LevelDesc& L, int k) {
    constexpr uint64_t sb = FMT == 1 ? 16u : 8u;
    TabLevel r;
    r.lvl = (const char*)vol + L.base * sb;
    r.tx = tr.lds + tr.t->x[k] + 1;
    r.ty = (const uint2*)(tr.lds + tr.t->y[k]) + 1;
    r.tz = (const uint2*)(tr.lds + tr.t->z[k]) + 1;
    return r;
}

__device__ __forceinline__ int step_level(const ConeDesc& cd, int i) {
    return (int)((cd.llev4[i >> 2] >> (8 * (i & 3))) & 0xffu);
}

// textureLod at level lv with fraction f toward lv + 1 (the step's lambda
// split on the host): (1-f)*tau1 + f*tau2.
// NCH 1: r only (the weighted R-sum cone); NCH 2: (r, a) of an RGBA volume,
// returned as RGBA (r, r, r, a).
// LA, LB: descriptors of levels lv and lv + 1 (LB unused when f == 0).
template <int FMT, int NCH>
__device__ __forceinline__ float4 sample_levels(const void* __restrict__ vol, const LevelDesc& LA, const LevelDesc& LB,
                                                float f, float sx, float sy, float sz, bool* two) {
    const bool tw = (__float_as_uint(f) << 1) != 0u;  // f != 0, as an integer test on the uniform f
    *two = tw;
    if constexpr (NCH == 2) {
        GreyTaps A, B;
        grey_fetch<FMT>(vol, LA, sx, sy, sz, A);
        if (tw) grey_fetch<FMT>(vol, LB, sx, sy, sz, B);  // in flight with A
        float2 a = grey_filter_f<FMT>(A);
        if (tw) {
            float2 b = grey_filter_f<FMT>(B);
            a.x = lerpF(a.x, b.x, f);
            a.y = lerpF(a.y, b.y, f);
        }
        return make_float4(a.x, a.x, a.x, a.y);
    } else {
        float r = trilinear_r<FMT>(vol, LA, sx, sy, sz);
        if (tw) r = lerpF(r, trilinear_r<FMT>(vol, LB, sx, sy, sz), f);
        return make_float4(r, 0.0f, 0.0f, 0.0f);
    }
}
template <int FMT, int NCH>
__device__ __forceinline__ float4 sample_lod(const void* __restrict__ vol, const VolumeDesc& vd, int lv, float f,
                                             float sx, float sy, float sz, bool* two) {
    return sample_levels<FMT, NCH>(vol, vd.lv[lv], vd.lv[lv + 1 < kMaxLevels ? lv + 1 : lv], f, sx, sy, sz, two);
}

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }
__device__ __forceinline__ uint32_t q8(float v) { return (uint32_t)(int)floorf(clamp01(v) * 255.0f + 0.5f); }

// Front-to-back cone state (modes 1/2) and march over steps [i0, i1):
//   stop before a sample outside [0,1]^3 (returns 1), accumulate
//   w = (1 - a) * alpha_i;  c += w * s.rgb;  a += w  (Kahan-compensated when
//   requested), stop once a >= alpha_stop (returns 2); returns 0 after i1.
// RGBA volumes are grey (r = g = b), so every sample has s.r = s.g = s.b and
// the three colour sums run the same operations on the same values: one sum
// c stands for all three, bit for bit.
struct FtbState {
    float c, a;
    float k, ka;  // compensation terms
};

// RGBA volumes are grey (r = g = b), so only r and a are filtered --
// bit-identical to filtering all four channels.  KAHAN: compensated sums.
// TAB: corner offsets from the workgroup's LDS tables (bit-identical loads).
template <int FMT, bool KAHAN, bool TAB>
__device__ __forceinline__ int march_ftb_t(const VolumeDesc& vd, const void* __restrict__ vol, const ConeDesc& cd,
                                           float p0x, float p0y, float p0z, float dx, float dy, float dz, int i0,
                                           int i1, FtbState& st, uint32_t& nsamp, uint32_t& ntwo, const TabRef& tr) {
    if (i0 >= i1) return 0;
    // a step's scalars are loaded one step ahead, its level descriptors only
    // when the level changes (at most once per level over a march)
    float h_n = cd.h[i0], f_n = cd.lfrac[i0];
    uint32_t lw_n = cd.llev4[i0 >> 2];  // the level's word, shifted at its step
    int cur = -1;
    LevelDesc LA{}, LB{};
    TabLevel TA{}, TB{};
    for (int i = i0; i < i1; ++i) {
        const float h = h_n, fr = f_n;
        const int lv = (int)((lw_n >> (8 * (i & 3))) & 0xffu);
        if (i + 1 < i1) {
            h_n = cd.h[i + 1];
            f_n = cd.lfrac[i + 1];
            lw_n = cd.llev4[(i + 1) >> 2];
        }
        if (lv != cur) {
            const int lb = lv + 1 < kMaxLevels ? lv + 1 : lv;
            LA = vd.lv[lv];
            LB = vd.lv[lb];
            if constexpr (TAB) {
                TA = tab_level<FMT>(tr, vol, LA, lv);
                TB = tab_level<FMT>(tr, vol, LB, lb);
            }
            cur = lv;
            // wait here (scalar loads complete out of order, so a later wait
            // for them would also wait for the step prefetch above)
            __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
        }
        float sx = fmaf(h, dx, p0x), sy = fmaf(h, dy, p0y), sz = fmaf(h, dz, p0z);
        if (!(sx >= 0.0f && sx <= 1.0f && sy >= 0.0f && sy <= 1.0f && sz >= 0.0f && sz <= 1.0f)) return 1;
        bool two;
        float4 sm;
        if constexpr (TAB) {
            // sample_levels<FMT, 2> with table addressing; f != 0 tested on
            // the scalar bits (fr is the same for every lane)
            uint32_t fr2;
            asm("s_lshl_b32 %0, %1, 1" : "=s"(fr2) : "s"(__builtin_amdgcn_readfirstlane(__float_as_uint(fr))));
            two = fr2 != 0u;
            GreyTaps A, B;
            grey_fetch_tab<FMT>(TA.lvl, TA.tx, TA.ty, TA.tz, LA.fn, sx, sy, sz, A);
            if (two) grey_fetch_tab<FMT>(TB.lvl, TB.tx, TB.ty, TB.tz, LB.fn, sx, sy, sz, B);
            float2 a = grey_filter_f<FMT>(A);
            if (two) {
                float2 b = grey_filter_f<FMT>(B);
                a.x = lerpF(a.x, b.x, fr);
                a.y = lerpF(a.y, b.y, fr);
            }
            sm = make_float4(a.x, a.x, a.x, a.y);
        } else {
            sm = sample_levels<FMT, 2>(vol, LA, LB, fr, sx, sy, sz, &two);
        }
        if (cd.count) {
            // counter runs only; the empty asm keeps this a scalar branch
            // (no per-step selects in the timed kernels)
            asm volatile("");
            ++nsamp;
            ntwo += two ? 1u : 0u;
        }
        float al;
        if (cd.cone_mode == 1) {
            al = sm.w;
        } else {
            float r = h * cd.tanHalf;
            float kk = r * cd.kscale;
            al = 1.0f - cv_exp2_neg(-(kk * sm.w));
        }
        float wgt = (1.0f - st.a) * al;
        if constexpr (KAHAN) {
            float y, t;
            y = fmaf(wgt, sm.x, -st.k); t = st.c + y; st.k = (t - st.c) - y; st.c = t;
            y = wgt - st.ka; t = st.a + y; st.ka = (t - st.a) - y; st.a = t;
        } else {
            st.c = fmaf(wgt, sm.x, st.c);
            st.a = st.a + wgt;
        }
        if (st.a >= cd.alpha_stop) return 2;
    }
    return 0;
}
template <int FMT, bool TAB = false>
__device__ __forceinline__ int march_ftb(const VolumeDesc& vd, const void* __restrict__ vol, const ConeDesc& cd,
                                         float p0x, float p0y, float p0z, float dx, float dy, float dz, int i0,
                                         int i1, FtbState& st, uint32_t& nsamp, uint32_t& ntwo,
                                         const TabRef& tr = TabRef{nullptr, nullptr}) {
    return cd.compensated
               ? march_ftb_t<FMT, true, TAB>(vd, vol, cd, p0x, p0y, p0z, dx, dy, dz, i0, i1, st, nsamp, ntwo, tr)
               : march_ftb_t<FMT, false, TAB>(vd, vol, cd, p0x, p0y, p0z, dx, dy, dz, i0, i1, st, nsamp, ntwo, tr);
}

// Far-field features after k samples: last sample position (2x - 1), cone
// direction (d * N), sun direction, and (density, light) at LODs 3..6 at
// distance max(h_k, feat_h[j]) along the cone, (0, 0) outside [0,1]^3.
template <int FMT>
__device__ __forceinline__ void farfield_features(const VolumeDesc& vd, const void* __restrict__ vol,
                                                  const ConeDesc& cd, float px, float py, float pz, float dx,
                                                  float dy, float dz, float fN, float f[kMlpIn]) {
    const float hk = cd.h[cd.learned_k - 1];
    const float xx = fmaf(hk, dx, px), xy = fmaf(hk, dy, py), xz = fmaf(hk, dz, pz);
    f[0] = 2.0f * xx - 1.0f;
    f[1] = 2.0f * xy - 1.0f;
    f[2] = 2.0f * xz - 1.0f;
    f[3] = dx * fN;
    f[4] = dy * fN;
    f[5] = dz * fN;
    f[6] = cd.sun[0];
    f[7] = cd.sun[1];
    f[8] = cd.sun[2];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float hj = fmaxf(hk, cd.feat_h[j]);
        const float sx = fmaf(hj, dx, px), sy = fmaf(hj, dy, py), sz = fmaf(hj, dz, pz);
        float dens = 0.0f, light = 0.0f;
        if (sx >= 0.0f && sx <= 1.0f && sy >= 0.0f && sy <= 1.0f && sz >= 0.0f && sz <= 1.0f) {
            GreyTaps T;
            grey_fetch<FMT>(vol, vd.lv[cd.feat_lvl[j]], sx, sy, sz, T);
            const float2 t = grey_filter_f<FMT>(T);
            dens = t.y;
            light = t.x;
        }
        f[9 + 2 * j] = dens;
        f[10 + 2 * j] = light;
        __builtin_amdgcn_sched_barrier(0);  // one footprint in flight at a time (bounds registers)
    }
}

// Cone start and direction of the fragment that billboard (cw: world centre
// and radius, vv: view-space centre) leaves at the pixel with view ray
// (rx, ry), where its sphere contribution sc = sphere_contrib(du, dv, s).
struct ConeFrag {
    float p0x, p0y, p0z, dx, dy, dz;
};
__device__ __forceinline__ ConeFrag cone_frag(const GridDesc& g, const CameraView& cv, const ConeDesc& cd, float4 cw,
                                              float du, float dv, float sc, float fN) {
    const float s = cw.w;
    ConeFrag f;
    // fragPos, then pos = fragPos + normalize(fragNor) * radius * sc
    float fx = (cw.x + du * cv.S[0]) + dv * cv.U[0];
    float fy = (cw.y + du * cv.S[1]) + dv * cv.U[1];
    float fz = (cw.z + du * cv.S[2]) + dv * cv.U[2];
    float qx = fx + (cv.B[0] * s) * sc;
    float qy = fy + (cv.B[1] * s) * sc;
    float qz = fz + (cv.B[2] * s) * sc;
    // calculateVoxelLerp(pos) / voxelDim
    f.p0x = div_N(g, fN * div_range(g, qx - g.lo[0], 0));
    f.p0y = div_N(g, fN * div_range(g, qy - g.lo[1], 1));
    f.p0z = div_N(g, fN * div_range(g, qz - g.lo[2], 2));
    if (cd.directional) {
        f.dx = cd.ldir[0]; f.dy = cd.ldir[1]; f.dz = cd.ldir[2];
    } else {
        float lx = cd.lpos[0] - qx, ly = cd.lpos[1] - qy, lz = cd.lpos[2] - qz;
        float inv = 1.0f / sqrtf((lx * lx + ly * ly) + lz * lz);
        f.dx = div_N(g, lx * inv); f.dy = div_N(g, ly * inv); f.dz = div_N(g, lz * inv);
    }
    return f;
}

// Blend one fragment's cone colour s into the pixel C: GL blend
// dst = src * src.a + dst * (1 - src.a) (src/main.cpp:94-95) for COMP = 0,
// front-to-back under for COMP = 1 (done once C.a >= alpha_stop).
template <int COMP>
__device__ __forceinline__ void blend_frag(const ConeDesc& cd, float s0, float s1, float s2, float s3, float& C0,
                                           float& C1, float& C2, float& C3, bool& done) {
    if (COMP == 0) {
        const float al = s3, om = 1.0f - al;
        C0 = s0 * al + C0 * om;
        C1 = s1 * al + C1 * om;
        C2 = s2 * al + C2 * om;
        C3 = s3 * al + C3 * om;
        if (cd.quantize) {
            C0 = (float)q8(C0) / 255.0f; C1 = (float)q8(C1) / 255.0f;
            C2 = (float)q8(C2) / 255.0f; C3 = (float)q8(C3) / 255.0f;
        }
    } else {
        const float wgt = (1.0f - C3) * s3;
        C0 = fmaf(wgt, s0, C0);
        C1 = fmaf(wgt, s1, C1);
        C2 = fmaf(wgt, s2, C2);
        C3 = C3 + wgt;
        if (C3 >= cd.alpha_stop) done = true;
    }
}

// ---------------------------------------------------------------------------
// Soft-particle noise march (res/conetrace_frag.glsl:103-120,136-174)
// ---------------------------------------------------------------------------
// floor(u) mod 32 as fl - 32 * floor(fl / 32): exact for every float fl (the
// scalings by 32 are exact and the difference is exact); NaN gives texel 0.
__device__ __forceinline__ int nm_wrap(float fl) {
    const float r = fl - 32.0f * floorf(fl * 0.03125f);
    return (r == r ? (int)r : 0) & 31;
}
// texture(noiseMap, s): GL_LINEAR, GL_REPEAT on the 32^3 map of decoded
// RGBA8_SNORM texels (x fastest); u = fma(s, 32, -1/2), alpha = u - floor(u),
// lerps along x, then y, then z
__device__ __forceinline__ float4 nm_tex(const float4* __restrict__ map, float sx, float sy, float sz) {
    const float ux = fmaf(sx, 32.0f, -0.5f), uy = fmaf(sy, 32.0f, -0.5f), uz = fmaf(sz, 32.0f, -0.5f);
    const float fx = floorf(ux), fy = floorf(uy), fz = floorf(uz);
    const float ax = ux - fx, ay = uy - fy, az = uz - fz;
    const int x0 = nm_wrap(fx), y0 = nm_wrap(fy), z0 = nm_wrap(fz);
    const int x1 = (x0 + 1) & 31, y1 = (y0 + 1) & 31, z1 = (z0 + 1) & 31;
    const int r00 = (z0 * 32 + y0) * 32, r10 = (z0 * 32 + y1) * 32, r01 = (z1 * 32 + y0) * 32, r11 = (z1 * 32 + y1) * 32;
    const float4 t000 = map[r00 + x0], t100 = map[r00 + x1], t010 = map[r10 + x0], t110 = map[r10 + x1];
    const float4 t001 = map[r01 + x0], t101 = map[r01 + x1], t011 = map[r11 + x0], t111 = map[r11 + x1];
    auto tri = [&](float a000, float a100, float a010, float a110, float a001, float a101, float a011, float a111) {
        const float c00 = lerpF(a000, a100, ax), c10 = lerpF(a010, a110, ax);
        const float c01 = lerpF(a001, a101, ax), c11 = lerpF(a011, a111, ax);
        return lerpF(lerpF(c00, c10, ay), lerpF(c01, c11, ay), az);
    };
    return make_float4(tri(t000.x, t100.x, t010.x, t110.x, t001.x, t101.x, t011.x, t111.x),
                       tri(t000.y, t100.y, t010.y, t110.y, t001.y, t101.y, t011.y, t111.y),
                       tri(t000.z, t100.z, t010.z, t110.z, t001.z, t101.z, t011.z, t111.z),
                       tri(t000.w, t100.w, t010.w, t110.w, t001.w, t101.w, t011.w, t111.w));
}
// glm dot: (x*x' + y*y') + z*z'
__device__ __forceinline__ float nm_dot(float ax, float ay, float az, float bx, float by, float bz) {
    return (ax * bx + ay * by) + az * bz;
}
// The noise branch of main() (res/conetrace_frag.glsl:136-174) for the
// fragment of billboard cw (centre, radius) at in-plane offset (du, dv):
// false where the view ray misses the sphere (disc < 0.01: discard); else col
// and runningOpacity * alpha.  noise3D (:103-120): octave i samples
// (uv + octaveOffsets[i]) * freq, the vec3 uniform indexed by component, 0 past
// the third; alpha = |sum|.  fragTex - 0.5 = (vertPos.xy + 1) / 2 - 0.5 with
// vertPos.xy = (du, dv) / radius.
__device__ __forceinline__ bool noise_fragment(const float4* __restrict__ map, const ConeDesc& cd, const CameraView& cv,
                                            float4 cw, float du, float dv, float& col_out, float& alpha_out) {
    const float r = cw.w;
    const float fpx = (cw.x + du * cv.S[0]) + dv * cv.U[0];
    const float fpy = (cw.y + du * cv.S[1]) + dv * cv.U[1];
    const float fpz = (cw.z + du * cv.S[2]) + dv * cv.U[2];
    const float rx = cd.nm_ray[0], ry = cd.nm_ray[1], rz = cd.nm_ray[2];
    const float ex = fpx - cw.x, ey = fpy - cw.y, ez = fpz - cw.z;
    const float A = nm_dot(rx, ry, rz, rx, ry, rz);
    const float B = 2.0f * nm_dot(ex, ey, ez, rx, ry, rz);
    const float C = nm_dot(ex, ey, ez, ex, ey, ez) - r * r;
    const float disc = B * B - (4.0f * A) * C;
    if (disc < 0.01f) return false;
    const float sq = sqrtf(disc);
    const float tnear = (-B - sq) / (2.0f * A);
    const float tfar = (-B + sq) / (2.0f * A);
    const float wnx = fpx + rx * tnear, wny = fpy + ry * tnear, wnz = fpz + rz * tnear;
    const float wfx = fpx + rx * tfar, wfy = fpy + ry * tfar, wfz = fpz + rz * tfar;
    float ux = (wnx - cw.x) / r, uy = (wny - cw.y) / r, uz = (wnz - cw.z) / r;
    const float fadj = 1.0f / cd.nm_adjust;
    float cx = wnx * fadj, cy = wny * fadj, cz = wnz * fadj;
    const float dlx = wfx * fadj - cx, dly = wfy * fadj - cy, dlz = wfz * fadj - cz;
    float steps = sqrtf(nm_dot(dlx, dly, dlz, dlx, dly, dlz)) / cd.nm_step;
    const float span = (float)(cd.nm_max_steps - cd.nm_min_steps);
    steps = (span < steps ? span : steps) + (float)cd.nm_min_steps;
    const float den = steps - 1.0f;
    const float dtx = dlx / den, dty = dly / den, dtz = dlz / den;
    const float opAdj = cd.nm_opacity / den;
    const float lightAdj = 1.0f / den;
    float ro = 0.0f, rl = 0.0f;
    for (int i = 0; (float)i < steps; ++i) {
        float n0 = 0.0f, n1 = 0.0f, n2 = 0.0f, n3 = 0.0f;
        float freq = 1.0f, pers = 1.0f;
        for (int o = 0; o < cd.nm_octaves; ++o) {
            const float off = o < 3 ? cd.nm_off[o] : 0.0f;
            const float4 t = nm_tex(map, (cx + off) * freq, (cy + off) * freq, (cz + off) * freq);
            n0 = n0 + pers * t.x; n1 = n1 + pers * t.y; n2 = n2 + pers * t.z; n3 = n3 + pers * t.w;
            freq = freq * cd.nm_freq_step;
            pers = pers * cd.nm_pers_step;
        }
        n3 = fabsf(n3);
        const float uu = nm_dot(ux, uy, uz, ux, uy, uz);
        const float inv = 1.0f / sqrtf(uu);
        const float nx = n0 + ux * inv, ny = n1 + uy * inv, nz = n2 + uz * inv;
        ro = ro + n3 * (1.0f - uu);
        rl = rl + clamp01(nm_dot(nx, ny, nz, 0.0f, 1.0f, 0.0f) * 0.5f + 0.5f);
        cx = cx + dtx; cy = cy + dty; cz = cz + dtz;
        ux = ux + dtx; uy = uy + dty; uz = uz + dtz;
    }
    col_out = cd.nm_min_color + (cd.nm_color_scale * rl) * lightAdj;
    const float tx = (du / r + 1.0f) / 2.0f - 0.5f, ty = (dv / r + 1.0f) / 2.0f - 0.5f;
    const float alpha = 1.0f - sqrtf(tx * tx + ty * ty) * 2.0f;
    alpha_out = clamp01(ro * opAdj) * alpha;
    return true;
}

// Ranks gathered per pass of the cone walk.
constexpr int kConeChunk = 256;

// CONE: 0 weighted R-sum, 1 front-to-back.  COMP: 0 GL blend, 1 front-to-back.
// LEARNED (front-to-back only): after cd.learned_k samples the far field of
// every still-running cone is predicted by the MLP (one 64-row batch per
// wave on the matrix cores); a cone whose predicted error, scaled by its
// remaining transmittance, exceeds cd.err_bound (or every cone when
// err_bound < 0) falls back to the plain march from step k.
//
// Shade one camera tile, one pixel per lane, each lane walking only the
// billboards that cover its own pixel.  The tile's binned billboards are
// taken in rank order (front to back for COMP = 1, back to front for
// COMP = 0) in chunks: the wave tests a chunk's billboards against its 64
// pixels and keeps those that cover any pixel, with the coverage ballot, in
// LDS (cover[], id[]); then every lane steps through that list to its own
// next covering billboard and all lanes march their own fragments together.
// A pixel composites exactly the fragments, in exactly the order, of a walk
// over every billboard, so the image is the same bit for bit; what changes is
// that a march no longer idles the lanes the billboard misses.  All lanes of
// a round start at step 0, so the LOD of step i (and with it the one- or
// two-level sampling branch) stays uniform across the wave, and in the
// learned kernels the lanes reach step k together: their cones form the
// wave's 64-row MLP batch.
template <int FMT, int CONE, int COMP, int LEARNED, int NM = 0, bool TAB = false>
__device__ __forceinline__ void cone_tile(const GridDesc& g, const VolumeDesc& vd, const CameraView& cv,
                                          const ConeDesc& cd, const void* __restrict__ vol,
                                          const float4* __restrict__ sph_world, const float4* __restrict__ sph_view,
                                          const int32_t* __restrict__ order,
                                          const unsigned long long* __restrict__ masks, int words,
                                          void* __restrict__ image, const MlpDev& mlp, uint16_t* xs,
                                          unsigned long long* cover, int* id, int tile, int lane, uint32_t& nfrag,
                                          uint32_t& nsamp, uint32_t& ntwo, uint32_t& nrows, uint32_t& nfall,
                                          uint32_t& nevals, const float4* __restrict__ nmap = nullptr,
                                          const TabRef& tr = TabRef{nullptr, nullptr}) {
    const int tx = tile % cv.tiles_x, ty = tile / cv.tiles_x;
    const int px = tx * cv.tile_w + (lane & (cv.tile_w - 1)), py = ty * cv.tile_h + (lane >> cv.tile_shift);
    const bool valid = px < cv.W && py < cv.H;
    const float rx = pixel_ndc(px, cv.W) * cv.kx;
    const float ry = pixel_ndc(py, cv.H) * cv.ky;
    const float fN = (float)g.N;
    const int kk_steps = LEARNED ? min(cd.learned_k, cd.steps) : cd.steps;

    float C0, C1, C2, C3;
    if (COMP == 0) {
        C0 = cd.clear[0]; C1 = cd.clear[1]; C2 = cd.clear[2]; C3 = cd.clear[3];
        if (cd.quantize) {
            C0 = (float)q8(C0) / 255.0f; C1 = (float)q8(C1) / 255.0f;
            C2 = (float)q8(C2) / 255.0f; C3 = (float)q8(C3) / 255.0f;
        }
        if (cd.sun_on && valid) {
            // the sun sprite, blended over the clear colour before the clouds
            // (res/sun_frag.glsl:14-28): fragment at the pixel ray's hit with
            // the sprite's view-aligned plane, dist = distance(center, fragPos)
            const float dz = -cd.sun_v[2];
            const float du = rx * dz - cd.sun_v[0], dv = ry * dz - cd.sun_v[1];
            if (fabsf(du) <= cd.sun_r1 && fabsf(dv) <= cd.sun_r1) {
                const float dist = sqrtf(du * du + dv * dv);
                float q0, q1, q2, q3;
                bool keep = true;
                if (dist < cd.sun_r0) {
                    q0 = cd.sun_in[0]; q1 = cd.sun_in[1]; q2 = cd.sun_in[2]; q3 = 1.0f;
                } else {
                    const float sc = (dist - cd.sun_r0) / (cd.sun_r1 - cd.sun_r0);
                    keep = !(sc > 0.99f);
                    q0 = cd.sun_out[0] * sc + cd.sun_in[0] * (1.0f - sc);
                    q1 = cd.sun_out[1] * sc + cd.sun_in[1] * (1.0f - sc);
                    q2 = cd.sun_out[2] * sc + cd.sun_in[2] * (1.0f - sc);
                    q3 = 1.0f - sc;
                }
                if (keep) {
                    bool unused = false;
                    blend_frag<0>(cd, clamp01(q0), clamp01(q1), clamp01(q2), clamp01(q3), C0, C1, C2, C3, unused);
                }
            }
        }
    } else {
        C0 = C1 = C2 = C3 = 0.0f;
    }
    bool done = !valid;
    const unsigned long long* tm = masks + (size_t)tile * words;
    int wi = 0, wcur = 0;
    unsigned long long bits = 0;

    while (!__all(done)) {
        // gather: the next kConeChunk binned billboards that cover any live pixel
        int n = 0;
        while (n < kConeChunk) {
            while (bits == 0 && wi < words) {
                wcur = COMP == 1 ? (words - 1 - wi) : wi;
                bits = tm[wcur];
                ++wi;
            }
            if (bits == 0) break;
            int b;
            if (COMP == 1) b = 63 - __clzll((long long)bits);
            else b = __ffsll((long long)bits) - 1;
            bits &= ~(1ull << b);
            const int j = order[(wcur << 6) | b];
            const float4 cw = sph_world[j];
            const float4 vv = sph_view[j];
            bool cov = false;
            if (!done) {
                const float dz = -vv.z;
                const float sc = sphere_contrib(rx * dz - vv.x, ry * dz - vv.y, cw.w);
                cov = !(sc < 0.01f);
            }
            const unsigned long long m = __ballot(cov);
            if (m) {
                if (lane == 0) {
                    cover[n] = m;
                    id[n] = j;
                }
                ++n;
            }
        }
        if (n == 0) break;
        __builtin_amdgcn_wave_barrier();
        // march: each lane its own next covering billboard, until none is left
        int cur = 0;
        for (;;) {
            bool have = false;
            int j = 0;
            if (!done) {
                while (cur < n) {
                    const unsigned long long m = cover[cur];
                    const int jj = id[cur];
                    ++cur;
                    if ((m >> lane) & 1ull) {
                        have = true;
                        j = jj;
                        break;
                    }
                }
            }
            if (!__any(have)) break;
            ConeFrag f = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            FtbState st = {0.0f, 0.0f, 0.0f, 0.0f};
            float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
            float s0u = 0.0f;  // weighted cone: the unclamped indirect
            float fr = 0.0f, fg = 0.0f, fb = 0.0f;  // colour with the predicted far field
            bool pred = false;
            bool alive = false;
            float nm_col = 1.0f, nm_alpha = 0.0f;  // noise march (NM): colour factor and fragment alpha
            if constexpr (NM) {
                // the noise branch runs first and may discard the fragment
                if (have) {
                    const float4 cw = sph_world[j];
                    const float4 vv = sph_view[j];
                    const float dz = -vv.z;
                    have = noise_fragment(nmap, cd, cv, cw, rx * dz - vv.x, ry * dz - vv.y, nm_col, nm_alpha);
                }
            }
            if (have) {
                const float4 cw = sph_world[j];
                const float4 vv = sph_view[j];
                const float dz = -vv.z;
                const float du = rx * dz - vv.x;
                const float dv = ry * dz - vv.y;
                const float sc = sphere_contrib(du, dv, cw.w);
                ++nfrag;
                f = cone_frag(g, cv, cd, cw, du, dv, sc, fN);
                if constexpr (CONE == 0) {
                    float color = 0.0f;
                    for (int i = 0; i < cd.steps; ++i) {
                        const float h = cd.h[i];
                        bool two;
                        float4 sm = sample_lod<FMT, 1>(vol, vd, step_level(cd, i), cd.lfrac[i], fmaf(h, f.dx, f.p0x), fmaf(h, f.dy, f.p0y),
                                                       fmaf(h, f.dz, f.p0z), &two);
                        ntwo += two ? 1u : 0u;
                        color = color + (sm.x * (float)(i + 1)) / cd.den;
                    }
                    nsamp += (uint32_t)cd.steps;
                    s0 = s1 = s2 = s3 = clamp01(color);
                    s0u = color;
                } else {
                    const int rs = march_ftb<FMT, TAB>(vd, vol, cd, f.p0x, f.p0y, f.p0z, f.dx, f.dy, f.dz, 0,
                                                             kk_steps, st, nsamp, ntwo, tr);
                    alive = LEARNED && rs == 0 && kk_steps < cd.steps;
                }
            }
            if constexpr (LEARNED && CONE != 0) {
                // wave-uniform: all lanes take part in the batched MLP
                if (__any(alive)) {
                    float x[kMlpIn];
                    if (alive) {
                        farfield_features<FMT>(vd, vol, cd, f.p0x, f.p0y, f.p0z, f.dx, f.dy, f.dz, fN, x);
                        nsamp += 4;
                    } else {
#pragma unroll
                        for (int q = 0; q < kMlpIn; ++q) x[q] = 0.0f;
                    }
                    uint32_t* xrow = (uint32_t*)(xs + lane * kMlpInPad);
#pragma unroll
                    for (int q = 0; q < kMlpInPad / 2; ++q) {
                        float lo = 2 * q < kMlpIn ? x[2 * q] : 0.0f;
                        float hi = 2 * q + 1 < kMlpIn ? x[2 * q + 1] : 0.0f;
                        xrow[q] = pack_bf2(lo, hi);
                    }
                    __syncthreads();
                    float y[kMlpOut];
                    mlp_forward_wave(mlp, xs, lane, y);
                    __syncthreads();
                    ++nevals;
                    if (alive) {
                        ++nrows;
                        const float rem = 1.0f - st.a;
                        if (cd.err_bound < 0.0f || !(rem * y[4] <= cd.err_bound)) {
                            ++nfall;
                            march_ftb<FMT>(vd, vol, cd, f.p0x, f.p0y, f.p0z, f.dx, f.dy, f.dz, kk_steps,
                                                 cd.steps, st, nsamp, ntwo);
                        } else {
                            fr = fmaf(rem, clamp01(y[0]), st.c);
                            fg = fmaf(rem, clamp01(y[1]), st.c);
                            fb = fmaf(rem, clamp01(y[2]), st.c);
                            st.a = st.a + rem * clamp01(y[3]);
                            pred = true;
                        }
                    }
                }
            }
            if (have) {
                if (CONE != 0) {
                    s0 = s1 = s2 = clamp01(st.c);
                    if (pred) { s0 = clamp01(fr); s1 = clamp01(fg); s2 = clamp01(fb); }
                    s3 = clamp01(st.a);
                }
                if constexpr (NM) {
                    // color = vec4(vec3(col), runningOpacity * alpha); color.rgb *= indirect
                    // (res/conetrace_frag.glsl:172, :194-196), clamped
                    const float ind = CONE != 0 ? st.c : s0u;
                    s0 = s1 = s2 = clamp01(nm_col * ind);
                    s3 = clamp01(nm_alpha);
                }
                blend_frag<COMP>(cd, s0, s1, s2, s3, C0, C1, C2, C3, done);
            }
        }
        // the next gather overwrites the list every lane has finished reading
        __builtin_amdgcn_wave_barrier();
    }
    if (valid) {
        char* row = (char*)image + (size_t)py * cd.pitch;
        if (cd.pixfmt == 0) {
            ((float4*)row)[px] = make_float4(C0, C1, C2, C3);
        } else {
            uint32_t v = q8(C0) | (q8(C1) << 8) | (q8(C2) << 16) | (q8(C3) << 24);
            ((uint32_t*)row)[px] = v;
        }
    }
}

// The learned variant asks for 3 waves per SIMD: the inline MLP needs ~100
// registers beside the march state, and without the bound the kernel gets 2.
template <int FMT, int CONE, int COMP, int LEARNED, int NM>
__global__ void __launch_bounds__(64, LEARNED ? 3 : (NM ? 4 : 6)) k_cone(GridDesc g, VolumeDesc vd, CameraView cv, ConeDesc cd,
                                             const void* __restrict__ vol, const float4* __restrict__ sph_world,
                                             const float4* __restrict__ sph_view, const int32_t* __restrict__ order,
                                             const unsigned long long* __restrict__ masks, int words,
                                             void* __restrict__ image, unsigned long long* __restrict__ counters,
                                             const int* __restrict__ tile_order, MlpDev mlp,
                                             const float4* __restrict__ nmap) {
    // One wavefront per tile; tiles are dispatched heaviest first
    // (tile_order), so the costly clump-centre tiles do not form the tail.
    __shared__ __align__(16) uint16_t xs[LEARNED ? 64 * kMlpInPad : 8];
    __shared__ unsigned long long cover[kConeChunk];
    __shared__ int id[kConeChunk];
    const int lane = threadIdx.x;
    uint32_t nfrag = 0, nsamp = 0, ntwo = 0, nrows = 0, nfall = 0, nevals = 0;
    const int tile = tile_order ? tile_order[blockIdx.x] : (int)blockIdx.x;
    cone_tile<FMT, CONE, COMP, LEARNED, NM>(g, vd, cv, cd, vol, sph_world, sph_view, order, masks, words, image,
                                            mlp, xs, cover, id, tile, lane, nfrag, nsamp, ntwo, nrows, nfall, nevals,
                                            nmap);
    if (counters) {
        unsigned long long a = wave_sum_u64(nfrag), bsum = wave_sum_u64(nsamp), c = wave_sum_u64(ntwo);
        unsigned long long e0 = 0, e1 = 0;
        if (LEARNED) { e0 = wave_sum_u64(nrows); e1 = wave_sum_u64(nfall); }
        if (lane == 0) {
            atomicAdd(&counters[3], a);
            atomicAdd(&counters[4], bsum);
            atomicAdd(&counters[5], c);
            if (LEARNED) {
                atomicAdd(&counters[6], e0);
                atomicAdd(&counters[7], e1);
                atomicAdd(&counters[8], (unsigned long long)nevals);
            }
        }
    }
}

// Build a workgroup's LDS tables (ConeTab): for level k of side n and
// j = floor(u) + 1 in [0, n], the byte offsets (from the level's first
// texel) of the clamped corners c0 = max(j - 1, 0) and c1 = min(j, n - 1):
// x of c0 (the slot holds x and x + 1), y and z of (c0, c1).
template <int FMT>
__device__ void build_cone_tabs(const VolumeDesc& vd, const ConeTab& t, uint32_t* lds) {
    constexpr uint32_t sb = FMT == 1 ? 16u : 8u;
    for (int k = 0; k < t.levels; ++k) {
        const LevelDesc L = vd.lv[k];
        const uint32_t n = L.n, srow = L.nb << 6, sslab = (L.nb * L.nb) << 6;
        uint32_t* tx = lds + t.x[k];
        uint2* ty = (uint2*)(lds + t.y[k]);
        uint2* tz = (uint2*)(lds + t.z[k]);
        for (uint32_t j = threadIdx.x; j <= n; j += blockDim.x) {
            const uint32_t c0 = j > 0u ? j - 1u : 0u, c1 = min(j, n - 1u);
            tx[j] = morton_off_x(c0) * sb;
            ty[j] = make_uint2(morton_off_y(c0, srow) * sb, morton_off_y(c1, srow) * sb);
            tz[j] = make_uint2(morton_off_z(c0, sslab) * sb, morton_off_z(c1, sslab) * sb);
        }
    }
    __syncthreads();
}

// The front-to-back cone trace with table addressing: t.wpg waves per
// workgroup share the LDS tables of every level (built once per
// workgroup); wave w of workgroup b shades tile tile_order[b * wpg + w].
// Dynamic LDS: the tables (t.words words), then per wave the gather lists
// (kConeChunk coverage masks and ids).
template <int FMT, int COMP>
__global__ void __launch_bounds__(512, 6) k_cone_tab(GridDesc g, VolumeDesc vd, CameraView cv, ConeDesc cd, ConeTab t,
                                                     const void* __restrict__ vol,
                                                     const float4* __restrict__ sph_world,
                                                     const float4* __restrict__ sph_view,
                                                     const int32_t* __restrict__ order,
                                                     const unsigned long long* __restrict__ masks, int words,
                                                     void* __restrict__ image, unsigned long long* __restrict__ counters,
                                                     const int* __restrict__ tile_order, int ntiles) {
    extern __shared__ __align__(16) uint32_t dyn[];
    build_cone_tabs<FMT>(vd, t, dyn);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int item = blockIdx.x * t.wpg + wave;
    if (item >= ntiles) return;
    unsigned long long* cover = (unsigned long long*)(dyn + t.words) + (size_t)wave * kConeChunk;
    int* id = (int*)((unsigned long long*)(dyn + t.words) + (size_t)t.wpg * kConeChunk) + (size_t)wave * kConeChunk;
    const int tile = tile_order ? tile_order[item] : item;
    uint32_t nfrag = 0, nsamp = 0, ntwo = 0, nrows = 0, nfall = 0, nevals = 0;
    const MlpDev none{nullptr, nullptr, nullptr};
    const TabRef tr{dyn, &t};
    cone_tile<FMT, 1, COMP, 0, 0, true>(g, vd, cv, cd, vol, sph_world, sph_view, order, masks, words, image, none,
                                        nullptr, cover, id, tile, lane, nfrag, nsamp, ntwo, nrows, nfall, nevals,
                                        nullptr, tr);
    if (counters) {
        unsigned long long a = wave_sum_u64(nfrag), b = wave_sum_u64(nsamp), c = wave_sum_u64(ntwo);
        if (lane == 0) {
            atomicAdd(&counters[3], a);
            atomicAdd(&counters[4], b);
            atomicAdd(&counters[5], c);
        }
    }
}

// Batched cone trace: global tiles of many scenes (heavy first), each wave
// finds its scene by binary search over the scenes' first tiles.
template <int FMT, int CONE, int COMP>
__global__ void __launch_bounds__(256) k_cone_batch(const BatchScene* __restrict__ scenes,
                                                    const int* __restrict__ tile0, int nscenes,
                                                    const int* __restrict__ gorder, int total,
                                                    unsigned long long* __restrict__ counters) {
    const int lane = threadIdx.x & 63;
    const int wpg = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    __shared__ unsigned long long cover[4][kConeChunk];  // waves per workgroup <= 4
    __shared__ int id[4][kConeChunk];
    uint32_t nfrag = 0, nsamp = 0, ntwo = 0, nrows = 0, nfall = 0, nevals = 0;
    const MlpDev none{nullptr, nullptr, nullptr};
    for (int item = blockIdx.x * wpg + wave; item < total; item += gridDim.x * wpg) {
        const int gt = __builtin_amdgcn_readfirstlane(gorder[item]);
        int lo = 0, hi = nscenes - 1;
        while (lo < hi) {  // last scene with tile0 <= gt
            const int mid = (lo + hi + 1) >> 1;
            if (tile0[mid] <= gt) lo = mid;
            else hi = mid - 1;
        }
        const int sidx = __builtin_amdgcn_readfirstlane(lo);
        const BatchScene& B = scenes[sidx];
        cone_tile<FMT, CONE, COMP, 0>(B.g, B.vd, B.cv, B.cd, B.levels, B.sph_world, B.sph_view, B.order, B.masks,
                                            B.words, B.image, none, nullptr, cover[wave], id[wave], gt - B.tile0, lane,
                                            nfrag, nsamp, ntwo, nrows, nfall, nevals);
    }
    if (counters) {
        unsigned long long a = wave_sum_u64(nfrag), b = wave_sum_u64(nsamp), c = wave_sum_u64(ntwo);
        if (lane == 0) {
            atomicAdd(&counters[3], a);
            atomicAdd(&counters[4], b);
            atomicAdd(&counters[5], c);
        }
    }
}

// Segment samples for training the far-field MLP: one thread per start point
// p0 (texture space), directional light.  Rows that are still running after
// k samples get the features, and as targets the rest of the plain march in
// units of the remaining transmittance: ((c - c_k) / (1 - a_k),
// (a - a_k) / (1 - a_k), 1 - a_k).
template <int FMT>
__global__ void __launch_bounds__(256) k_segments(GridDesc g, VolumeDesc vd, ConeDesc cd, const void* __restrict__ vol,
                                                  const float* __restrict__ p0, int n, float* __restrict__ feat,
                                                  float* __restrict__ tgt, int* __restrict__ valid) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float px = p0[3 * i], py = p0[3 * i + 1], pz = p0[3 * i + 2];
    const float dx = cd.ldir[0], dy = cd.ldir[1], dz = cd.ldir[2];
    const int kk_steps = cd.learned_k;
    FtbState st = {0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t ns = 0, nt = 0;
    int rs = march_ftb<FMT>(vd, vol, cd, px, py, pz, dx, dy, dz, 0, min(kk_steps, cd.steps), st, ns, nt);
    const bool rec = rs == 0 && kk_steps < cd.steps;
    float f[kMlpIn];
    float ck = 0.0f, ak = 0.0f;
    if (rec) {
        farfield_features<FMT>(vd, vol, cd, px, py, pz, dx, dy, dz, (float)g.N, f);
        ck = st.c; ak = st.a;
        march_ftb<FMT>(vd, vol, cd, px, py, pz, dx, dy, dz, kk_steps, cd.steps, st, ns, nt);
    }
    valid[i] = rec ? 1 : 0;
#pragma unroll
    for (int q = 0; q < kMlpIn; ++q) feat[(size_t)i * kMlpIn + q] = rec ? f[q] : 0.0f;
    float* t = tgt + (size_t)i * kMlpOut;
    if (rec) {
        const float om = 1.0f - ak;
        t[0] = t[1] = t[2] = (st.c - ck) / om;  // grey
        t[3] = (st.a - ak) / om;
        t[4] = om;
    } else {
        t[0] = t[1] = t[2] = t[3] = t[4] = 0.0f;
    }
}

// ---------------------------------------------------------------------------
// occupancy count/hash, state readback, level de-bricking
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_occupancy(GridDesc g, const unsigned long long* __restrict__ shellbits,
                                                   const unsigned long long* __restrict__ fillbits, uint32_t nbricks,
                                                   unsigned long long* __restrict__ out2) {
    uint32_t brick = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long cnt = 0, hash = 0;
    if (brick < nbricks) {
        unsigned long long sh = shellbits[brick];
        unsigned long long fl = fillbits ? fillbits[brick] : 0ull;
        unsigned long long flm = 0;  // fill bits in Morton order
        while (fl) {
            const uint32_t l = (uint32_t)(__ffsll((long long)fl) - 1);
            fl &= fl - 1ull;
            flm |= 1ull << morton6(l & 3u, (l >> 2) & 3u, l >> 4);
        }
        unsigned long long any = sh | flm;
        uint32_t nb = g.nb0;
        uint32_t bx = brick % nb, by = (brick / nb) % nb, bz = brick / (nb * nb);
        while (any) {
            int m = __ffsll((long long)any) - 1;
            any &= ~(1ull << m);
            uint32_t lx, ly, lz;
            demorton6((uint32_t)m, &lx, &ly, &lz);
            unsigned long long x = bx * 4u + lx, y = by * 4u + ly, z = bz * 4u + lz;
            unsigned long long st = ((sh >> m) & 1ull) ? 2ull : 1ull;
            unsigned long long lin = x + (unsigned long long)g.N * (y + (unsigned long long)g.N * z);
            unsigned long long zz = lin * 4ull + st;
            zz = (zz ^ (zz >> 30)) * 0xbf58476d1ce4e5b9ull;
            zz = (zz ^ (zz >> 27)) * 0x94d049bb133111ebull;
            zz ^= zz >> 31;
            hash += zz;
            ++cnt;
        }
    }
    cnt = wave_sum_u64(cnt);
    hash = wave_sum_u64(hash);
    if ((threadIdx.x & 63) == 0 && cnt) {
        atomicAdd(&out2[0], cnt);
        atomicAdd(&out2[1], hash);
    }
}

__global__ void __launch_bounds__(256) k_read_state(GridDesc g, const unsigned long long* __restrict__ shellbits,
                                                    const unsigned long long* __restrict__ fillbits,
                                                    uint8_t* __restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t N = (uint64_t)g.N;
    if (i >= N * N * N) return;
    uint32_t x = (uint32_t)(i % N), y = (uint32_t)((i / N) % N), z = (uint32_t)(i / (N * N));
    uint32_t w = ((z >> 2) * g.nb0 + (y >> 2)) * g.nb0 + (x >> 2);
    uint32_t m = morton6(x & 3u, y & 3u, z & 3u);
    uint8_t st = 0;
    if ((shellbits[w] >> m) & 1ull) st = 2;
    else if (fillbits && ((fillbits[w] >> lin6(x & 3u, y & 3u, z & 3u)) & 1ull)) st = 1;
    out[i] = st;
}

__global__ void __launch_bounds__(256) k_debrick(LevelDesc L, int tb, const uint8_t* __restrict__ levels,
                                                 uint8_t* __restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t n = L.n;
    if (i >= n * n * n) return;
    uint32_t x = (uint32_t)(i % n), y = (uint32_t)((i / n) % n), z = (uint32_t)(i / (n * n));
    uint64_t a = brick_addr(L, x, y, z);
    if (tb == 16) {  // slot (r, a, r+, a+) -> RGBA (r, r, r, a)
        const float2 t = ((const float2*)levels)[2 * a];
        ((float4*)out)[i] = make_float4(t.x, t.x, t.x, t.y);
    } else if (tb == 8) {  // halves (r, a, r+, a+) -> (r, r, r, a)
        const uint32_t u = ((const uint32_t*)levels)[2 * a];
        const uint32_t rr = u & 0xffffu;
        ((uint2*)out)[i] = make_uint2(rr | (rr << 16), rr | (u & 0xffff0000u));
    } else {
        out[i] = levels[a];
    }
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(uint64_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

size_t texel_bytes(int format) { return format == 0 ? 1 : (format == 1 ? 16 : 8); }

void launch_fill_multi(const FillSeg* segs, int nseg, hipStream_t s) {
    FillArgs a{};
    uint64_t total = 0;
    int k = 0;
    for (int i = 0; i < nseg && k < 4; ++i) {
        if (segs[i].n == 0 || !segs[i].p) continue;
        a.p[k] = segs[i].p;
        a.n[k] = segs[i].n;
        a.v[k] = segs[i].value;
        total += segs[i].n;
        ++k;
    }
    a.nseg = k;
    if (total == 0) return;
    for (int i = k; i < 4; ++i) a.p[i] = a.p[0];  // empty segments (n = 0)
    unsigned nb = blocks_for(total, 256);
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_fill_multi, dim3(nb), dim3(256), 0, s, a);
}

void launch_fill_u64(uint64_t* p, uint64_t n, uint64_t value, hipStream_t s) {
    if (n == 0) return;
    unsigned nb = blocks_for(n, 256);
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_fill_u64, dim3(nb), dim3(256), 0, s, p, n, value);
}


void launch_light_items(const GridDesc& g, const LightView& lv, int mode, const float4* boards, int S,
                        const LightScratch& ls, hipStream_t s) {
    if (S <= 0) return;
    const int cap = (int)std::min<size_t>(ls.items_cap, 0x7fffffff);
    const dim3 bgrid((unsigned)((S + 255) / 256));
    hipLaunchKernelGGL(k_light_boxes, bgrid, dim3(256), 0, s, g, lv, boards, S, ls.lbox, ls.counts, mode == 1 ? 1 : 0);
    hipLaunchKernelGGL(k_light_scan, dim3(1), dim3(256), 0, s, ls.counts, S, ls.offsets);
    hipLaunchKernelGGL(k_light_expand, bgrid, dim3(256), 0, s, ls.offsets, S, ls.items, cap);
    const dim3 igrid((unsigned)std::min<size_t>((ls.items_cap + 255) / 256, 2048));
    hipLaunchKernelGGL(k_light_prep, igrid, dim3(256), 0, s, g, lv, boards, ls.lbox, ls.offsets + S, ls.items, cap,
                       mode == 1 ? 1 : 0, raster_words(g));
}

void launch_light_raster(const GridDesc& g, const LightView& lv, int mode, const float4* boards, int S,
                         const LightScratch& ls, unsigned long long* keys, unsigned long long* fillbits,
                         unsigned long long* counters, hipStream_t s) {
    if (S <= 0) return;
    const int cap = (int)std::min<size_t>(ls.items_cap, 0x7fffffff);
    unsigned grid = (unsigned)std::min<size_t>(ls.items_cap, 4096);
    const bool p2 = g.pow2_range && g.pow2_N && g.pow2_step;
    const int chord = mode == 1 ? 1 : 0;
    if (p2 && raster_words(g) == 2048)
        hipLaunchKernelGGL((k_light_raster<true, 2048>), dim3(grid), dim3(256), 0, s, g, lv, chord, ls.offsets + S,
                           ls.items, cap, keys, fillbits, counters);
    else if (p2)
        hipLaunchKernelGGL((k_light_raster<true, 4096>), dim3(grid), dim3(256), 0, s, g, lv, chord, ls.offsets + S,
                           ls.items, cap, keys, fillbits, counters);
    else
        hipLaunchKernelGGL((k_light_raster<false, 4096>), dim3(grid), dim3(256), 0, s, g, lv, chord, ls.offsets + S,
                           ls.items, cap, keys, fillbits, counters);
}

size_t light_items_bound(int S, float max_radius, const LightView& lv, const GridDesc& g, int chord) {
    // depth slices multiply the items; a billboard's slice count grows with its radius
    const size_t kmax = raster_flat(g, chord) ? (size_t)light_slices_box(g, lv, max_radius, raster_words(g)) : 1;
    // widest pixel box of a billboard of radius max_radius: the box spans
    // s * W / lsx pixels plus at most 5 from floor/ceil and the 1-pixel margins
    double wx = 2.0 * max_radius * lv.W / (2.0 * lv.lsx) + 6.0;
    double wy = 2.0 * max_radius * lv.H / (2.0 * lv.lsy) + 6.0;
    size_t tx = std::min<size_t>((size_t)(wx / kLightTile) + 2, (size_t)(lv.W / kLightTile) + 1);
    size_t ty = std::min<size_t>((size_t)(wy / kLightTile) + 2, (size_t)(lv.H / kLightTile) + 1);
    return std::max<size_t>((size_t)S * tx * ty * kmax, 1);
}

void launch_light_splat(const GridDesc& g, const LightView& lv, const float4* boards, const unsigned long long* keys,
                        float4* posmap, unsigned long long* shellbits, unsigned long long* counters, hipStream_t s) {
    const uint64_t tiles = (uint64_t)((lv.W + 15) / 16) * (uint64_t)((lv.H + 15) / 16);
    hipLaunchKernelGGL(k_light_splat, dim3((unsigned)tiles), dim3(256), 0, s, g, lv, boards, keys, posmap, shellbits,
                       counters);
}

// level 0 and the whole mip chain; levels 0..4 in one fused pass when the
// side is a multiple of 16, the remaining levels one launch each
size_t build_zero_flags(const GridDesc& g, const VolumeDesc& vd) {
    if (g.N % 16 != 0 || vd.format == 0) return 0;
    const size_t rpa = (size_t)g.N / 16;
    return rpa * rpa * rpa;
}

void launch_build_levels(const GridDesc& g, const VolumeDesc& vd, const unsigned long long* shellbits,
                         const unsigned long long* fillbits, void* levels, uint8_t* zflags, hipStream_t s) {
    int first_generic = 1;
    if (g.N % 16 == 0) {
        const uint32_t rpa = (uint32_t)g.N / 16u;
        const int nlev = vd.L < 5 ? vd.L : 5;
        const dim3 grid(rpa * rpa * rpa);
        const bool eval = vd.mode == 1 && vd.noise_octaves > 0 && !vd.ntex;  // noise without its texture
        if (vd.format == 0)
            hipLaunchKernelGGL(k_build_fused_r8, grid, dim3(256), 0, s, g, vd, shellbits, levels, nlev, rpa);
        else if (vd.format == 1 && eval)
            hipLaunchKernelGGL((k_build_fused<1, false>), grid, dim3(256), 0, s, g, vd, shellbits, fillbits, levels,
                               zflags, nlev, rpa);
        else if (vd.format == 1)
            hipLaunchKernelGGL((k_build_fused<1, true>), grid, dim3(256), 0, s, g, vd, shellbits, fillbits, levels,
                               zflags, nlev, rpa);
        else if (eval)
            hipLaunchKernelGGL((k_build_fused<2, false>), grid, dim3(256), 0, s, g, vd, shellbits, fillbits, levels,
                               zflags, nlev, rpa);
        else
            hipLaunchKernelGGL((k_build_fused<2, true>), grid, dim3(256), 0, s, g, vd, shellbits, fillbits, levels,
                               zflags, nlev, rpa);
        first_generic = nlev;
    } else {
        uint32_t nbricks = g.nb0 * g.nb0 * g.nb0;
        hipLaunchKernelGGL(k_level0, dim3(blocks_for(nbricks, 4)), dim3(256), 0, s, g, vd, shellbits, fillbits, levels,
                           nbricks);
    }
    int k = first_generic;
    for (; k < vd.L && vd.lv[k].n * vd.lv[k].n * vd.lv[k].n > (uint32_t)kTailTexels; ++k) launch_mip(vd, k, levels, s);
    if (k < vd.L) hipLaunchKernelGGL(k_mip_tail, dim3(1), dim3(1024), 0, s, vd, k, levels);
}

void launch_mip(const VolumeDesc& vd, int k, void* levels, hipStream_t s) {
    uint32_t nb = vd.lv[k].nb;
    uint32_t ntex = nb * nb * nb * 64u;
    hipLaunchKernelGGL(k_mip, dim3(blocks_for(ntex, 256)), dim3(256), 0, s, vd, k, levels, ntex);
}

void launch_cam_setup(const GridDesc& g, const CameraView& cv, const float4* boards, int S, float4* sph_world,
                      float4* sph_view, int4* sph_tiles, unsigned long long* keys, hipStream_t s) {
    if (S <= 0) return;
    hipLaunchKernelGGL(k_cam_setup, dim3(blocks_for(S, 256)), dim3(256), 0, s, g, cv, boards, S, sph_world, sph_view,
                       sph_tiles, keys);
}

// Scratch `keys` (sort_scratch_words(S) u64 words): [0, npad) the keys written
// by k_cam_setup; for S beyond one LDS: [npad, 2 npad) padded keys, then
// idx u32[npad], tie bits u32[npad / 32], pos i32[npad], at i32[npad].
size_t sort_scratch_words(int S) {
    size_t npad = 2;
    while (npad < (size_t)S) npad <<= 1;
    return 5 * npad;
}

template <int E>
static int launch_sort_lds(const unsigned long long* keys, int S, int32_t* order, int* tie_flag, hipStream_t s) {
    static_assert(1024 * E <= kLdsSortMax, "one workgroup's LDS");
    const int bytes = 1024 * E * (int)sizeof(unsigned long long);
    if (bytes > 64 * 1024) {
        static bool attr_set = false;  // raise the dynamic LDS limit once per instantiation
        if (!attr_set) {
            if (hipFuncSetAttribute((const void*)k_sort_lds<E>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) !=
                hipSuccess)
                return -1;
            attr_set = true;
        }
    }
    hipLaunchKernelGGL(k_sort_lds<E>, dim3(1), dim3(1024), (size_t)bytes, s, keys, S, order, tie_flag);
    return 1024 * E;
}

void launch_sort_keys(const GridDesc& g, const float4* boards, int S, const float point[3], unsigned long long* keys,
                      hipStream_t s) {
    if (S <= 0) return;
    hipLaunchKernelGGL(k_sort_keys, dim3(blocks_for(S, 256)), dim3(256), 0, s, g, boards, S, point[0], point[1],
                       point[2], keys);
}

void launch_iota(int32_t* order, int S, hipStream_t s) {
    if (S <= 0) return;
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(S, 256)), dim3(256), 0, s, order, S);
}

int launch_sort(unsigned long long* keys, int S, int32_t* order, int* tie_flag, hipStream_t s) {
    if (S <= 0) return 0;
    int npad = 2;
    while (npad < S) npad <<= 1;
    if (npad <= kLdsSortMax) {
        int r;
        switch (npad < 1024 ? 1024 : npad) {
            case 1024: r = launch_sort_lds<1>(keys, S, order, tie_flag, s); break;
            case 2048: r = launch_sort_lds<2>(keys, S, order, tie_flag, s); break;
            case 4096: r = launch_sort_lds<4>(keys, S, order, tie_flag, s); break;
            case 8192: r = launch_sort_lds<8>(keys, S, order, tie_flag, s); break;
            default: r = launch_sort_lds<16>(keys, S, order, tie_flag, s); break;
        }
        if (r > 0) return r;  // else (LDS limit not raised): the global-memory sort below
    }
    unsigned long long* pk = keys + npad;
    uint32_t* g_idx = (uint32_t*)(keys + 2 * (size_t)npad);
    uint32_t* g_tie = g_idx + npad;
    int32_t* pos = (int32_t*)(g_tie + npad / 32);
    int32_t* at = pos + npad;
    hipLaunchKernelGGL(k_pad_keys, dim3(blocks_for(npad, 256)), dim3(256), 0, s, keys, pk, S, npad);
    for (int k = 2; k <= npad; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1)
            hipLaunchKernelGGL(k_bitonic_global, dim3(blocks_for(npad, 256)), dim3(256), 0, s, pk, npad, j, k);
    launch_fill_u64((uint64_t*)tie_flag, 1, 0, s);  // the flag buffer holds 64 bytes
    launch_fill_u64((uint64_t*)g_tie, (uint64_t)npad / 64, 0, s);
    hipLaunchKernelGGL(k_sort_finish, dim3(blocks_for(S, 256)), dim3(256), 0, s, pk, S, order, g_idx, g_tie, tie_flag);
    hipLaunchKernelGGL(k_sort_replay_global, dim3(1), dim3(1), 0, s, S, g_idx, g_tie, pos, at, order, tie_flag);
    return npad;
}

size_t tile_scratch_ints(int ntiles) { return 66 + 3 * (size_t)ntiles + 2; }

void launch_bin_counts(const int4* sph_tiles, const int32_t* order, int S, int tiles_x, int words,
                       unsigned long long* masks, int* tile_count, hipStream_t s) {
    if (S > 0)
        hipLaunchKernelGGL(k_bin, dim3(blocks_for(S, 4)), dim3(256), 0, s, sph_tiles, order, S, tiles_x, words, masks,
                           tile_count);
}

// heavy-first order of ntiles tile counts; scratch: hist[32] | cursor[32] | bucket_of[ntiles]
void launch_tile_order(const int* tile_count, int ntiles, int* scratch, int* tile_order, hipStream_t s) {
    int* hist = scratch;
    int* cursor = scratch + 32;
    int* bucket_of = scratch + 64;
    launch_fill_u64((uint64_t*)scratch, 32, 0, s);
    hipLaunchKernelGGL(k_tile_bucket, dim3(blocks_for(ntiles, 256)), dim3(256), 0, s, tile_count, ntiles, bucket_of,
                       hist);
    hipLaunchKernelGGL(k_tile_scatter, dim3(blocks_for(ntiles, 256)), dim3(256), 0, s, bucket_of, ntiles, hist, cursor,
                       tile_order);
}

// Binning plus the heavy-first tile order.  tile scratch (ints): hist[32] |
// cursor[32] | (2 unused) | tile_count[ntiles] | bucket_of | tile_order.  The
// masks and the scratch up to tile_count are cleared by one fill launch.
TileQueue launch_bin(const int4* sph_tiles, const int32_t* order, int S, int tiles_x, int tiles_y, int words,
                     unsigned long long* masks, int* scratch, hipStream_t s) {
    const int ntiles = tiles_x * tiles_y;
    unsigned int* counter = (unsigned int*)(scratch + 64);
    int* tile_count = scratch + 66;
    int* tile_order = tile_count + 2 * (size_t)ntiles;
    const FillSeg segs[2] = {{(uint64_t*)masks, (uint64_t)ntiles * (uint64_t)words, 0},
                             {(uint64_t*)scratch, (uint64_t)(66 + ntiles + 1) / 2, 0}};
    launch_fill_multi(segs, 2, s);
    launch_bin_counts(sph_tiles, order, S, tiles_x, words, masks, tile_count, s);
    int* hist = scratch;
    int* cursor = scratch + 32;
    int* bucket_of = tile_count + ntiles;
    hipLaunchKernelGGL(k_tile_bucket, dim3(blocks_for(ntiles, 256)), dim3(256), 0, s, tile_count, ntiles, bucket_of,
                       hist);
    hipLaunchKernelGGL(k_tile_scatter, dim3(blocks_for(ntiles, 256)), dim3(256), 0, s, bucket_of, ntiles, hist, cursor,
                       tile_order);
    TileQueue q;
    q.counter = counter;
    q.order = tile_order;
    return q;
}

template <int FMT, int CONE, int COMP>
static void launch_cone_batch_t(const BatchScene* scenes, const int* tile0, int nscenes, const int* order, int total,
                                int wpg, int groups, unsigned long long* counters, hipStream_t s) {
    if (wpg < 1 || wpg > 4) return;  // k_cone_batch holds chunk lists for at most 4 waves
    const int items_per_group = wpg;
    unsigned grid = groups > 0 ? (unsigned)groups : (unsigned)((total + items_per_group - 1) / items_per_group);
    if (grid == 0) return;
    hipLaunchKernelGGL((k_cone_batch<FMT, CONE, COMP>), dim3(grid), dim3(64 * wpg), 0, s, scenes, tile0, nscenes,
                       order, total, counters);
}

void launch_cone_batch(const BatchScene* scenes, const int* tile0, int nscenes, const int* order, int total_tiles,
                       int format, int cone_mode, int composite, int waves_per_group, int groups,
                       unsigned long long* counters, hipStream_t s) {
    if (total_tiles <= 0) return;
    const bool w = cone_mode == 0;
    const int c = composite == 0 ? 0 : 1;
    auto go = [&](auto fmt) {
        constexpr int F = decltype(fmt)::value;
        if (w && c == 0) launch_cone_batch_t<F, 0, 0>(scenes, tile0, nscenes, order, total_tiles, waves_per_group, groups, counters, s);
        else if (w) launch_cone_batch_t<F, 0, 1>(scenes, tile0, nscenes, order, total_tiles, waves_per_group, groups, counters, s);
        else if (c == 0) launch_cone_batch_t<F, 1, 0>(scenes, tile0, nscenes, order, total_tiles, waves_per_group, groups, counters, s);
        else launch_cone_batch_t<F, 1, 1>(scenes, tile0, nscenes, order, total_tiles, waves_per_group, groups, counters, s);
    };
    if (format == 1) go(std::integral_constant<int, 1>{});
    else go(std::integral_constant<int, 2>{});
}

// dynamic LDS of k_cone_tab: the tables and, per wave, kConeChunk coverage
// masks (8 bytes) and billboard ids (4 bytes)
static size_t cone_tab_lds_bytes_for(const ConeTab& t, int wpg) {
    return (size_t)t.words * 4u + (size_t)wpg * kConeChunk * (8u + 4u);
}
size_t cone_tab_lds_bytes(const ConeTab& t) { return cone_tab_lds_bytes_for(t, t.wpg); }

bool cone_tab_plan(const VolumeDesc& vd, ConeTab& t) {
    t = ConeTab{};
    if (vd.format != 1 && vd.format != 2) return false;
    const uint64_t sb = vd.format == 1 ? 16u : 8u;
    uint32_t w = 0;
    t.levels = vd.L;
    for (int k = 0; k < vd.L; ++k) {
        const LevelDesc& L = vd.lv[k];
        const uint64_t bytes = (uint64_t)L.nb * L.nb * L.nb * 64u * sb;
        if (bytes >= (1ull << 32)) return false;
        const uint32_t n1 = L.n + 1u;
        w = (w + 1u) & ~1u;  // the pair tables are read as uint2
        t.y[k] = w; w += 2u * n1;
        t.z[k] = w; w += 2u * n1;
        t.x[k] = w; w += n1;
    }
    t.words = (w + 3u) & ~3u;  // the gather lists that follow are 16-byte aligned
    // workgroups of 4 waves while four of them fit beside the VGPR-bound 6
    // waves per SIMD, else 8 (the tables are shared by more waves)
    t.wpg = cone_tab_lds_bytes_for(t, 4) * 6 <= 160u * 1024u ? 4 : 8;
    return cone_tab_lds_bytes(t) <= 64u * 1024u;
}

// The noise-march kernels (NM = 1) are separate instantiations, so the plain
// kernels keep their register budget.
template <int FMT, int CONE, int COMP, int LEARNED>
static void launch_cone_t(const ConeLaunch& a, hipStream_t s) {
    unsigned ntiles = (unsigned)(a.cv.tiles_x * a.cv.tiles_y);
    if constexpr (!LEARNED && CONE == 1 && FMT != 0) {
        ConeTab t;
        if (!a.nmap && a.table_addressing && cone_tab_plan(a.vd, t)) {
            const unsigned grid = (ntiles + (unsigned)t.wpg - 1u) / (unsigned)t.wpg;
            hipLaunchKernelGGL((k_cone_tab<FMT, COMP>), dim3(grid), dim3(64 * t.wpg), cone_tab_lds_bytes(t), s, a.g,
                               a.vd, a.cv, a.cd, t, a.levels, a.sph_world, a.sph_view, a.order, a.masks, a.words,
                               a.image, a.counters, a.tile_order, (int)ntiles);
            return;
        }
    }
    if constexpr (!LEARNED) {
        if (a.nmap) {
            hipLaunchKernelGGL((k_cone<FMT, CONE, COMP, 0, 1>), dim3(ntiles), dim3(64), 0, s, a.g, a.vd, a.cv, a.cd,
                               a.levels, a.sph_world, a.sph_view, a.order, a.masks, a.words, a.image, a.counters,
                               a.tile_order, a.mlp, a.nmap);
            return;
        }
    }
    hipLaunchKernelGGL((k_cone<FMT, CONE, COMP, LEARNED, 0>), dim3(ntiles), dim3(64), 0, s, a.g, a.vd, a.cv, a.cd,
                       a.levels, a.sph_world, a.sph_view, a.order, a.masks, a.words, a.image, a.counters,
                       a.tile_order, a.mlp, a.nmap);
}

// R8 volumes hold one channel: only the weighted R-sum cone reads them (the
// runtime rejects the front-to-back modes for R8 before launching).
template <int FMT>
static void launch_cone_f(const ConeLaunch& a, hipStream_t s) {
    const bool weighted = a.cd.cone_mode == 0;
    if constexpr (FMT == 0) {
        if (a.cd.composite == 0) launch_cone_t<0, 0, 0, 0>(a, s);
        else launch_cone_t<0, 0, 1, 0>(a, s);
    } else {
        if (a.learned) launch_cone_t<FMT, 1, 1, 1>(a, s);
        else if (weighted && a.cd.composite == 0) launch_cone_t<FMT, 0, 0, 0>(a, s);
        else if (weighted) launch_cone_t<FMT, 0, 1, 0>(a, s);
        else if (a.cd.composite == 0) launch_cone_t<FMT, 1, 0, 0>(a, s);
        else launch_cone_t<FMT, 1, 1, 0>(a, s);
    }
}

void launch_cone(const ConeLaunch& a, hipStream_t s) {
    if (a.vd.format == 0) launch_cone_f<0>(a, s);
    else if (a.vd.format == 1) launch_cone_f<1>(a, s);
    else launch_cone_f<2>(a, s);
}

void launch_segments(const GridDesc& g, const VolumeDesc& vd, const ConeDesc& cd, const void* levels,
                     const float* p0, int n, float* feat, float* tgt, int* valid, hipStream_t s) {
    if (n <= 0) return;
    if (vd.format == 1)
        hipLaunchKernelGGL(k_segments<1>, dim3(blocks_for((uint64_t)n, 256)), dim3(256), 0, s, g, vd, cd, levels, p0,
                           n, feat, tgt, valid);
    else
        hipLaunchKernelGGL(k_segments<2>, dim3(blocks_for((uint64_t)n, 256)), dim3(256), 0, s, g, vd, cd, levels, p0,
                           n, feat, tgt, valid);
}

void launch_occupancy(const GridDesc& g, const unsigned long long* shellbits, const unsigned long long* fillbits,
                      unsigned long long* out2, hipStream_t s) {
    uint32_t nbricks = g.nb0 * g.nb0 * g.nb0;
    hipLaunchKernelGGL(k_occupancy, dim3(blocks_for(nbricks, 256)), dim3(256), 0, s, g, shellbits, fillbits, nbricks,
                       out2);
}

void launch_read_state(const GridDesc& g, const unsigned long long* shellbits, const unsigned long long* fillbits,
                       uint8_t* out, hipStream_t s) {
    uint64_t n = (uint64_t)g.N * g.N * g.N;
    hipLaunchKernelGGL(k_read_state, dim3(blocks_for(n, 256)), dim3(256), 0, s, g, shellbits, fillbits, out);
}

void launch_debrick(const VolumeDesc& vd, int k, const void* levels, void* out, hipStream_t s) {
    uint64_t n = (uint64_t)vd.lv[k].n;
    n = n * n * n;
    hipLaunchKernelGGL(k_debrick, dim3(blocks_for(n, 256)), dim3(256), 0, s, vd.lv[k], (int)texel_bytes(vd.format),
                       (const uint8_t*)levels, (uint8_t*)out);
}


// ---------------------------------------------------------------------------
// Batched stage kernels: one launch per stage over the scenes of a batch.
// Workgroup b finds its group member m by binary search over the stage's
// first-workgroup table and runs the single-scene kernel's body on scene
// scene[m] with its own workgroup index b - first[m] and count
// first[m + 1] - first[m]: the same code and the same workgroup indices as
// the scene's own launch, so every output is the same.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int stage_member(const BatchStage& st, int b) {
    int lo = 0, hi = st.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (st.first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}
struct StageSlot {
    int sc, bid, nblk;
};
__device__ __forceinline__ StageSlot stage_slot(const BatchStage& st) {
    const int m = stage_member(st, (int)blockIdx.x);
    StageSlot r;
    r.sc = __builtin_amdgcn_readfirstlane(st.scene[m]);
    r.bid = (int)blockIdx.x - __builtin_amdgcn_readfirstlane(st.first[m]);
    r.nblk = __builtin_amdgcn_readfirstlane(st.first[m + 1]) - __builtin_amdgcn_readfirstlane(st.first[m]);
    return r;
}

__global__ void __launch_bounds__(256) kb_fill(const FillSeg* __restrict__ segs, BatchStage st) {
    const StageSlot q = stage_slot(st);
    const FillSeg f = segs[q.sc];
    const uint64_t e0 = (uint64_t)q.bid * kFillChunk, e1 = min(f.n, e0 + kFillChunk);
    for (uint64_t e = e0 + threadIdx.x; e < e1; e += 256) f.p[e] = f.value;
}
__global__ void __launch_bounds__(256) kb_light_boxes(const BatchVox* __restrict__ sc, BatchStage st) {
    const StageSlot q = stage_slot(st);
    const BatchVox& B = sc[q.sc];
    k_light_boxes_body(B.g, B.lv, B.boards, B.S, B.lbox, B.counts, B.chord, q.bid, q.nblk);
}
__global__ void __launch_bounds__(256) kb_light_scan(const BatchVox* __restrict__ sc, BatchStage st) {
    const StageSlot q = stage_slot(st);
    const BatchVox& B = sc[q.sc];
    k_light_scan_body(B.counts, B.S, B.offsets);
}
__global__ void __launch_bounds__(256) kb_light_expand(const BatchVox* __restrict__ sc, BatchStage st) {
    const StageSlot q = stage_slot(st);
    const BatchVox& B = sc[q.sc];
    k_light_expand_body(B.offsets, B.S, B.items, (int)min(B.items_cap, 0x7fffffff), q.bid, q.nblk);
}
__global__ void __launch_bounds__(256) kb_light_prep(const BatchVox* __restrict__ sc, BatchStage st) {
    const StageSlot q = stage_slot(st);
    const BatchVox& B = sc[q.sc];
    k_light_prep_body(B.g, B.lv, B.boards, B.lbox, B.offsets + B.S, B.items, (int)min(B.items_cap, 0x7fffffff), B.chord,
                      B.words, q.bid, q.nblk);
}
template <bool P2, int WORDS>
__global__ void __launch_bounds__(256) kb_light_raster(const BatchVox* __restrict__ sc, BatchStage st) {
    const StageSlot q = stage_slot(st);
    const BatchVox& B = sc[q.sc];
    k_light_raster_body<P2, WORDS>(B.g, B.lv, B.chord, B.offsets + B.S, B.items, (int)min(B.items_cap, 0x7fffffff),
                                   B.keys, B.fill, nullptr, q.bid, q.nblk);
}
__global__ void __launch_bounds__(256) kb_light_splat(const BatchVox* __restrict__ sc, BatchStage st) {
    const StageSlot q = stage_slot(st);
    const BatchVox& B = sc[q.sc];
    k_light_splat_body(B.g, B.lv, B.boards, B.keys, nullptr, B.shell, nullptr, q.bid, q.nblk);
}
template <int FMT, bool TEX>
__global__ void __launch_bounds__(256) kb_build_fused(const BatchVox* __restrict__ sc, BatchStage st) {
    const StageSlot q = stage_slot(st);
    const BatchVox& B = sc[q.sc];
    k_build_fused_body<FMT, TEX>(B.g, B.vd, B.shell, B.vd.mode == 1 ? B.fill : nullptr, B.levels, B.zflags, B.nlev,
                                 B.rpa, q.bid, q.nblk);
}
__global__ void __launch_bounds__(1024) kb_mip_tail(const BatchVox* __restrict__ sc, BatchStage st) {
    const StageSlot q = stage_slot(st);
    const BatchVox& B = sc[q.sc];
    k_mip_tail_body(B.vd, B.tail_k, B.levels);
}
__global__ void __launch_bounds__(256) kb_cam_setup(const BatchVox* __restrict__ sc, BatchStage st) {
    const StageSlot q = stage_slot(st);
    const BatchVox& B = sc[q.sc];
    k_cam_setup_body(B.g, B.cv, B.boards, B.S, B.sph_world, B.sph_view, B.sph_tiles, B.sortkeys, q.bid, q.nblk);
}
__global__ void __launch_bounds__(256) kb_iota(const BatchVox* __restrict__ sc, BatchStage st) {
    const StageSlot q = stage_slot(st);
    const BatchVox& B = sc[q.sc];
    const int j = q.bid * 256 + (int)threadIdx.x;
    if (j < B.S) B.order[j] = j;
}
template <int E>
__global__ void __launch_bounds__(1024) kb_sort_lds(const BatchVox* __restrict__ sc, BatchStage st) {
    const StageSlot q = stage_slot(st);
    const BatchVox& B = sc[q.sc];
    k_sort_lds_body<E>(B.sortkeys, B.S, B.order, B.tie);
}
__global__ void __launch_bounds__(256) kb_bin(const BatchVox* __restrict__ sc, BatchStage st) {
    const StageSlot q = stage_slot(st);
    const BatchVox& B = sc[q.sc];
    k_bin_body(B.sph_tiles, B.order, B.S, B.cv.tiles_x, B.mask_words, B.masks, B.tile_count, q.bid, q.nblk);
}

int fill_segment_blocks(uint64_t n) { return (int)((n + kFillChunk - 1) / kFillChunk); }

void launch_fill_segments(const FillSeg* d_segs, const int* d_first, int nseg, int total_blocks, hipStream_t s) {
    if (total_blocks <= 0) return;
    const BatchStage st{d_first, d_first + nseg + 1, nseg};
    hipLaunchKernelGGL(kb_fill, dim3((unsigned)total_blocks), dim3(256), 0, s, d_segs, st);
}

bool batch_vox_eligible(const GridDesc& g, const VolumeDesc& vd, int S) {
    if (vd.format != 1 && vd.format != 2) return false;
    if (g.N % 16 != 0 || S > kBatchSortMax) return false;
    for (int k = 5; k < vd.L; ++k)
        if (vd.lv[k].n * vd.lv[k].n * vd.lv[k].n > (uint32_t)kTailTexels) return false;
    return true;
}

static int sort_width(int S) {
    int npad = 1024;
    while (npad < S) npad <<= 1;
    return npad / 1024;
}

void batch_stage_shape(int stage, const BatchVox& b, int* blocks, int* variant) {
    const int S = b.S;
    *variant = 0;
    *blocks = 0;
    switch (stage) {
        case kBsBoxes:
        case kBsExpand:
        case kBsCam:
            *blocks = S > 0 ? (S + 255) / 256 : 0;
            break;
        case kBsScan:
            *blocks = S > 0 ? 1 : 0;
            break;
        case kBsPrep:
            *blocks = S > 0 ? (int)std::min<size_t>(((size_t)b.items_cap + 255) / 256, 2048) : 0;
            break;
        case kBsRaster: {
            *blocks = S > 0 ? (int)std::min<size_t>((size_t)b.items_cap, 4096) : 0;
            const bool p2 = b.g.pow2_range && b.g.pow2_N && b.g.pow2_step;
            *variant = p2 ? (raster_words(b.g) == 2048 ? 0 : 1) : 2;
            break;
        }
        case kBsSplat:
            *blocks = (int)(((b.lv.W + 15) / 16) * ((b.lv.H + 15) / 16));
            break;
        case kBsBuild:
            *blocks = (int)(b.rpa * b.rpa * b.rpa);
            *variant = b.vd.format | (b.vd.mode == 1 && b.vd.noise_octaves > 0 && !b.vd.ntex ? 4 : 0);
            break;
        case kBsTail:
            *blocks = b.tail_k < b.vd.L ? 1 : 0;
            break;
        case kBsSortIota:
            *blocks = (S > 0 && b.order_mode == 1) ? (S + 255) / 256 : 0;
            break;
        case kBsSort:
            *blocks = (S > 0 && b.order_mode != 1) ? 1 : 0;
            *variant = sort_width(S);
            break;
        case kBsBin:
            *blocks = S > 0 ? (S + 3) / 4 : 0;
            break;
        default:
            break;
    }
}

void batch_vox_complete(BatchVox& b) {
    b.words = raster_words(b.g);
    b.rpa = (uint32_t)b.g.N / 16u;
    b.nlev = b.vd.L < 5 ? b.vd.L : 5;
    b.tail_k = b.nlev;
}

void launch_batch_stage(int stage, int variant, const BatchVox* scenes, const BatchStage& st, int total_blocks,
                        hipStream_t s) {
    if (total_blocks <= 0) return;
    const dim3 grid((unsigned)total_blocks);
    switch (stage) {
        case kBsBoxes: hipLaunchKernelGGL(kb_light_boxes, grid, dim3(256), 0, s, scenes, st); break;
        case kBsScan: hipLaunchKernelGGL(kb_light_scan, grid, dim3(256), 0, s, scenes, st); break;
        case kBsExpand: hipLaunchKernelGGL(kb_light_expand, grid, dim3(256), 0, s, scenes, st); break;
        case kBsPrep: hipLaunchKernelGGL(kb_light_prep, grid, dim3(256), 0, s, scenes, st); break;
        case kBsRaster:
            if (variant == 0) hipLaunchKernelGGL((kb_light_raster<true, 2048>), grid, dim3(256), 0, s, scenes, st);
            else if (variant == 1) hipLaunchKernelGGL((kb_light_raster<true, 4096>), grid, dim3(256), 0, s, scenes, st);
            else hipLaunchKernelGGL((kb_light_raster<false, 4096>), grid, dim3(256), 0, s, scenes, st);
            break;
        case kBsSplat: hipLaunchKernelGGL(kb_light_splat, grid, dim3(256), 0, s, scenes, st); break;
        case kBsBuild:
            // variant: format | 4 when the noise is evaluated (no texture)
            if (variant == 1) hipLaunchKernelGGL((kb_build_fused<1, true>), grid, dim3(256), 0, s, scenes, st);
            else if (variant == 5) hipLaunchKernelGGL((kb_build_fused<1, false>), grid, dim3(256), 0, s, scenes, st);
            else if (variant == 6) hipLaunchKernelGGL((kb_build_fused<2, false>), grid, dim3(256), 0, s, scenes, st);
            else hipLaunchKernelGGL((kb_build_fused<2, true>), grid, dim3(256), 0, s, scenes, st);
            break;
        case kBsTail: hipLaunchKernelGGL(kb_mip_tail, grid, dim3(1024), 0, s, scenes, st); break;
        case kBsCam: hipLaunchKernelGGL(kb_cam_setup, grid, dim3(256), 0, s, scenes, st); break;
        case kBsSortIota: hipLaunchKernelGGL(kb_iota, grid, dim3(256), 0, s, scenes, st); break;
        case kBsSort: {
            const size_t bytes = (size_t)1024 * variant * sizeof(unsigned long long);
            if (variant == 1) hipLaunchKernelGGL(kb_sort_lds<1>, grid, dim3(1024), bytes, s, scenes, st);
            else if (variant == 2) hipLaunchKernelGGL(kb_sort_lds<2>, grid, dim3(1024), bytes, s, scenes, st);
            else if (variant == 4) hipLaunchKernelGGL(kb_sort_lds<4>, grid, dim3(1024), bytes, s, scenes, st);
            else hipLaunchKernelGGL(kb_sort_lds<8>, grid, dim3(1024), bytes, s, scenes, st);
            break;
        }
        case kBsBin: hipLaunchKernelGGL(kb_bin, grid, dim3(256), 0, s, scenes, st); break;
        default: break;
    }
}

}  // namespace cv
